"""Many independent PDHG instances advanced together, one HIP-graph launch per iteration.

The reference reconstructs one problem per invocation and launches every kernel
from Python in every iteration (``tvm.py:390-491``).  Here a set of independent,
unevenly sized instances (each a HIP solver with its own plan, grids and duals)
is advanced in lock step: one iteration of every instance still running is
captured once into a HIP graph -- each instance on its own stream, forked from
and joined back to a capture stream, so the runtime may run small instances
side by side -- and the graph is replayed once per iteration.  After each replay
the float64 stopping sums of all instances are read with one synchronisation;
an instance that meets its own rule and tolerance (or its iteration cap) leaves
the set and the graph is captured again for the rest.  Each instance runs the
same kernels in the same order as when solved alone, so its result is
bit-identical to ``solver.run()``.

mode="streams" issues the same per-instance iteration eagerly, each instance on
its own stream (so independent instances can share the GPU through its hardware
queues), and reads all stopping sums after one synchronisation per iteration.
"""

from __future__ import annotations

import ctypes
import time

from . import _lib


class GraphBatch:
    def __init__(self, solvers):
        import torch
        if not solvers:
            raise ValueError("no instances")
        self.solvers = list(solvers)
        self.device = self.solvers[0].device
        if any(s.device != self.device for s in self.solvers):
            raise ValueError("all instances must live on one GPU")
        self.origin = torch.cuda.Stream(self.device)
        self.streams = [torch.cuda.Stream(self.device) for _ in self.solvers]
        self.captures = 0

    def _prepare(self):
        """Run one eager iteration of every instance (creates every workspace the iteration
        touches, so the captured launches allocate nothing), then reset each to its start."""
        import torch
        for s in self.solvers:
            with torch.cuda.device(self.device):
                s.pdhg_init(None)
                s.reset_dual()
                s.pdhg(0)
                s.pdhg_init(None)
                s.reset_dual()
        torch.cuda.synchronize(self.device)

    def _capture(self, active):
        import torch
        lib = _lib.load()
        h = self.origin.cuda_stream
        torch.cuda.synchronize(self.device)
        graph = ctypes.c_void_p()
        with torch.cuda.device(self.device), torch.cuda.stream(self.origin):
            _lib.call("tr_graph_begin", h)
            try:
                for i in active:
                    st = self.streams[i]
                    st.wait_stream(self.origin)
                    with torch.cuda.stream(st):
                        self.solvers[i].pdhg_launch(0)
                    self.origin.wait_stream(st)
            finally:
                rc = lib.tr_graph_end(h, ctypes.byref(graph))
        _lib.check(rc, "tr_graph_end")
        self.captures += 1
        return graph

    def _launch_streams(self, active):
        import torch
        with torch.cuda.device(self.device):
            for i in active:
                st = self.streams[i]
                st.wait_stream(self.origin)
                with torch.cuda.stream(st):
                    self.solvers[i].pdhg_launch(0)
                self.origin.wait_stream(st)

    def run(self, on_replay=None, mode="graph"):
        """Advance every instance to its stopping rule.  Returns per-instance dicts with the
        iteration count, the seconds until it finished (from the first replay) and its history;
        each solver's ``ubar`` holds its result.  on_replay(k, active) runs after replay k."""
        import torch
        if mode not in ("graph", "streams"):
            raise ValueError(f"unknown mode {mode!r}")
        if mode == "graph" and any(getattr(s, "_bstreams", None) for s in self.solvers):
            # per-bin streams forked inside a capture are not supported: set_bin_streams(1) first
            raise ValueError("graph mode needs every instance on one bin stream (set_bin_streams(1))")
        self._prepare()
        n = len(self.solvers)
        its = [0] * n
        hist = [[] for _ in range(n)]
        done_at = [None] * n
        active = list(range(n))
        graph = self._capture(active) if mode == "graph" else None
        lib = _lib.load()
        t0 = time.perf_counter()
        k = 0
        try:
            while active:
                if graph is not None:
                    _lib.call("tr_graph_launch", graph, self.origin.cuda_stream)
                else:
                    self._launch_streams(active)
                self.origin.synchronize()
                sums = torch.cat([self.solvers[i]._red for i in active]).tolist()
                k += 1
                if on_replay is not None:
                    on_replay(k, list(active))
                finished = []
                for j, i in enumerate(active):
                    s = self.solvers[i]
                    its[i] += 1
                    tol = s._converged_value(sums[2 * j], sums[2 * j + 1])
                    hist[i].append(tol)
                    if tol < s.tol or its[i] >= s.max_iter:
                        finished.append(i)
                        done_at[i] = time.perf_counter() - t0
                if finished:
                    active = [i for i in active if i not in finished]
                    if graph is not None:
                        lib.tr_graph_destroy(graph)
                        graph = None
                        if active:
                            graph = self._capture(active)
        finally:
            if graph is not None:
                lib.tr_graph_destroy(graph)
        for i, s in enumerate(self.solvers):
            s.iterations = its[i]
            s.history = hist[i]
            s.elapsed = done_at[i]
        return [{"iterations": its[i], "seconds": done_at[i], "history": hist[i]} for i in range(n)]
