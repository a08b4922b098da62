"""Write PyWavelets db6 zero-mode golden vectors for the TVMW wavelet oracle.

Run with an interpreter that has PyWavelets (here /opt/conda/bin/python3.9 with
PyWavelets 1.1.1, a third-party library, not the reference):

    /opt/conda/bin/python3.9 tests/golden/make_pywt_golden.py

Writes tests/golden/pywt_db6_zero.npz holding the db6 filter bank, the
``wavedecn(..., 'db6', mode='zero', axes=(-3,-2,-1))`` coefficients of seeded
complex arrays at auto level (``dwtn_max_level``), and ``waverecn`` of those
coefficients.  The reference applies ptwt ``wavedec3``/``waverec3`` with the
same wavelet, mode and level rule (src/toporecon/algorithms/tvmw.py:196-206).
"""

import os

import numpy as np
import pywt

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pywt_db6_zero.npz")


def main():
    w = pywt.Wavelet("db6")
    data = {
        "rec_lo": np.array(w.rec_lo),
        "dec_lo": np.array(w.dec_lo),
        "dec_hi": np.array(w.dec_hi),
        "rec_hi": np.array(w.rec_hi),
    }
    # (22, 23, 24): level 1 with an odd axis; (44, 45, 46): level 2 with odd
    # intermediate lengths.  Inputs are float32-valued (exact in float64) to
    # keep the fixture small; the transform is real-linear, so real inputs
    # pin the complex transform too.
    shapes = [(22, 23, 24), (44, 45, 46)]
    rng = np.random.RandomState(1234)
    for i, shp in enumerate(shapes):
        x = rng.randn(*shp).astype(np.float32)
        level = pywt.dwtn_max_level(shp, "db6")
        coeffs = pywt.wavedecn(x.astype(np.float64), "db6", mode="zero", level=level,
                               axes=(-3, -2, -1))
        data[f"x{i}"] = x
        data[f"level{i}"] = np.array(level)
        data[f"cA{i}"] = coeffs[0]
        for lv, det in enumerate(coeffs[1:]):
            for key, val in det.items():
                data[f"d{i}_{lv}_{key}"] = val
        if i == 0:
            # waverecn of a second, independent coefficient set.
            c2 = [rng.randn(*coeffs[0].shape)] + [
                {k: rng.randn(*v.shape) for k, v in d.items()} for d in coeffs[1:]]
            data["rc_cA"] = c2[0]
            for lv, det in enumerate(c2[1:]):
                for key, val in det.items():
                    data[f"rc_d_{lv}_{key}"] = val
            data["rc_rec"] = pywt.waverecn(c2, "db6", mode="zero", axes=(-3, -2, -1))
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
