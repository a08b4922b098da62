"""The HIP-graph batch's capture sequence is well formed (CPU, recording stand-ins for torch's
stream layer and the library).

During a stream capture every launch must be issued on a stream that has joined the capture (it
waited on work of a stream already in it), and before the capture ends the origin stream must
depend on every launch.  The batch forks each instance's stream from the origin and joins it back;
an instance with several bin streams forks them from its own stream twice per iteration (data dual
update, data adjoint) and joins them back each time.  A recording of the waits (torch's
``wait_stream``: record an event on the waited stream, wait on it) and launches is replayed here
with the capture rules: the sequence passes for one and for several bin streams per instance, and
a sequence with a join left out is caught.
"""

import contextlib
import sys
import types

import pytest


class Recorder:
    def __init__(self):
        self.log = []
        self.n = 0


class FakeStream:
    def __init__(self, rec, name):
        self.rec = rec
        rec.n += 1
        self.cuda_stream = 1000 + rec.n
        self.name = name

    def wait_stream(self, other):
        self.rec.log.append(("wait", self.cuda_stream, other.cuda_stream))

    def synchronize(self):
        pass


def fake_torch(rec):
    cuda = types.SimpleNamespace()
    default = FakeStream(rec, "default")
    state = {"cur": default}

    def Stream(device=None):
        return FakeStream(rec, "side")

    def current_stream(device=None):
        return state["cur"]

    @contextlib.contextmanager
    def stream(s):
        prev = state["cur"]
        state["cur"] = s
        try:
            yield
        finally:
            state["cur"] = prev

    @contextlib.contextmanager
    def device(d):
        yield

    cuda.Stream = Stream
    cuda.current_stream = current_stream
    cuda.stream = stream
    cuda.device = device
    cuda.synchronize = lambda d=None: None
    torch = types.ModuleType("torch")
    torch.cuda = cuda
    return torch


class FakeSolver:
    """The stream behaviour of a solver's pdhg_launch: its own launches on the current stream and
    the per-bin data-term launches through the solver base's fork / join helpers."""

    def __init__(self, rec, bins, bin_streams):
        from toporecon_hip.algorithms._base import HipPdhgBase
        self.rec = rec
        self.B = bins
        self.device = 0
        self._bin_streams = HipPdhgBase._bin_streams.__get__(self)
        self._join_bin_streams = HipPdhgBase._join_bin_streams.__get__(self)
        self._bstreams = [FakeStream(rec, "bin") for _ in range(bin_streams)] if bin_streams > 1 else None
        self.drop_join = False

    def _launch(self, handle):
        self.rec.log.append(("launch", handle))

    def pdhg_launch(self, it):
        from toporecon_hip import _lib
        for phase in ("dual", "adjoint"):
            s = _lib.stream_handle(self.device)
            hs = self._bin_streams()
            for b in range(self.B):
                self._launch(hs[b] if hs else s)
            if not self.drop_join:
                self._join_bin_streams()
            self._launch(_lib.stream_handle(self.device))  # TV dual / primal + reduction


def replay_capture(log):
    """Apply the capture rules to a recorded sequence; returns the launches the origin misses."""
    begin = [i for i, e in enumerate(log) if e[0] == "begin"]
    assert len(begin) == 1
    origin = log[begin[0]][1]
    in_capture = {origin}
    covered = {origin: set()}
    work = set()
    for e in log[begin[0] + 1:]:
        if e[0] == "wait":
            _, waiter, waited = e
            if waited in in_capture:
                in_capture.add(waiter)
                covered[waiter] = covered.get(waiter, set()) | covered[waited]
        elif e[0] == "launch":
            s = e[1]
            assert s in in_capture, f"launch on stream {s} outside the capture"
            item = ("w", len(work))
            work.add(item)
            covered[s] = covered.get(s, set()) | {item}
        elif e[0] == "end":
            assert e[1] == origin
            return work - covered[origin]
    raise AssertionError("capture never ended")


def run_capture(monkeypatch, bin_streams, drop_join=False):
    rec = Recorder()
    torch = fake_torch(rec)
    monkeypatch.setitem(sys.modules, "torch", torch)
    from toporecon_hip import _lib, batch

    def call(name, *args):
        if name == "tr_graph_begin":
            rec.log.append(("begin", args[0]))
        return 0

    class Lib:
        @staticmethod
        def tr_graph_end(h, out):
            rec.log.append(("end", h))
            return 0

    monkeypatch.setattr(batch._lib, "call", call)
    monkeypatch.setattr(batch._lib, "load", lambda: Lib)
    monkeypatch.setattr(batch._lib, "check", lambda rc, what="": None)
    solvers = [FakeSolver(rec, bins, bin_streams) for bins in (3, 5, 2)]
    for s in solvers:
        s.drop_join = drop_join
    assert _lib.stream_handle(0) == torch.cuda.current_stream().cuda_stream
    b = batch.GraphBatch(solvers)
    b._capture([0, 1, 2])
    return rec.log


@pytest.mark.parametrize("bin_streams", [1, 3])
def test_capture_sequence_joins_every_launch(monkeypatch, bin_streams):
    log = run_capture(monkeypatch, bin_streams)
    launches = [e for e in log if e[0] == "launch"]
    assert len(launches) == 2 * (3 + 5 + 2) + 2 * 3
    assert replay_capture(log) == set()


def test_replay_catches_a_missing_join(monkeypatch):
    log = run_capture(monkeypatch, 3, drop_join=True)
    assert replay_capture(log)
