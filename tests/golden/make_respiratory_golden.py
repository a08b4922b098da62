"""Write tests/golden/respiratory_vectors.npz: input/output vectors of the reference's
respiratory self-gating, ``estimate_respiratory_signal`` in
src/toporecon/preprocessing/respiratory.py:36-114.

Nothing else pins that function to the bit: the reference's tests do not cover it, and its result
depends on the filter order, the float32 casts and the sign LAPACK's SVD gives the principal
component, none of which a closed form fixes.  Its only dependencies (NumPy, SciPy) are installed
here, so this script loads the reference's module file directly (the preprocessing package's
``__init__`` would also import the SigPy-based JSENSE, which is absent) and records its outputs on
seeded synthetic k-space centres.  The fixture holds data only: inputs, settings and outputs.

Cases: 3 echoes x 4 coils x 2000 trajectories at TR = 5 ms (both principal-component stages),
the first echo alone (the single-echo path), and the 3-echo input with a custom band
(0.2-1.5 Hz, orders 6 and 3).  Regenerate with ``python tests/golden/make_respiratory_golden.py``
in a container that holds the reference at /root/reference.
"""

import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = "/root/reference/src"


def synthetic_centres(seed: int = 7, echoes: int = 3, coils: int = 4, traj: int = 2000, tr: float = 0.005):
    """k-space centre samples whose magnitude carries a 0.27 Hz breathing trace, a slow drift and
    noise, with coil- and echo-dependent gains and random phases; shape (E, C, T, 1) complex64."""
    rng = np.random.default_rng(seed)
    t = np.arange(traj) * tr
    breath = np.sin(2 * np.pi * 0.27 * t + 0.4) + 0.3 * np.sin(2 * np.pi * 0.54 * t + 1.1)
    drift = 0.2 * t / t[-1]
    out = np.empty((echoes, coils, traj, 1), np.complex64)
    for e in range(echoes):
        for c in range(coils):
            gain = (1.0 + 0.3 * c) * np.exp(-0.25 * e)
            mag = gain * (5.0 + 0.4 * rng.uniform(0.5, 1.5) * breath + drift) + 0.05 * rng.standard_normal(traj)
            phase = rng.uniform(-np.pi, np.pi, traj)
            out[e, c, :, 0] = (mag * np.exp(1j * phase)).astype(np.complex64)
    return out


def main():
    sys.path.insert(0, REF_SRC)
    spec = importlib.util.spec_from_file_location(
        "reference_respiratory", os.path.join(REF_SRC, "toporecon", "preprocessing", "respiratory.py"))
    ref = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = ref  # dataclasses look their module up while the class is built
    spec.loader.exec_module(ref)
    tr = 0.005
    centres = synthetic_centres(tr=tr)
    custom = dict(low_cut_hz=0.2, high_cut_hz=1.5, low_pass_order=6, high_pass_order=3)
    out = dict(
        centres=centres,
        tr_seconds=np.float64(tr),
        resp_3echo=ref.estimate_respiratory_signal(centres, tr),
        resp_1echo=ref.estimate_respiratory_signal(centres[:1], tr),
        resp_custom=ref.estimate_respiratory_signal(centres, tr, config=ref.RespiratoryConfig(**custom)),
        custom=np.array([custom["low_cut_hz"], custom["high_cut_hz"], custom["low_pass_order"],
                         custom["high_pass_order"]], np.float64),
    )
    path = os.path.join(HERE, "respiratory_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
