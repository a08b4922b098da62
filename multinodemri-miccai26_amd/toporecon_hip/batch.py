"""Many independent PDHG instances advanced together in HIP graphs.

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

mode="persistent" captures one graph, once, for the whole solve: `chunk`
iterations of every instance, each iteration followed by the instance's stopping
rule evaluated on the device (``tr_stop_rule``), every launch of an instance gated
by its own device flag (``tr_gate_set``).  The rule closes the gate when the
instance meets its tolerance or its iteration cap, and every later kernel of that
instance returns at once, so its state stays exactly where a solve alone would
have stopped.  The graph is replayed unchanged until every gate is closed (one
host synchronisation per replay; with chunk >= the largest iteration cap a single
launch runs the whole batch); no instance waits for another between replays, so
small instances run beside large ones.  Iteration counts and histories come from
the device; per-instance times from device clock stamps taken when each gate
closes.
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

    def _capture_persistent(self, chunk, gates, its, hist, stamps, rules, maxit):
        import torch
        lib = _lib.load()
        h = self.origin.cuda_stream
        torch.cuda.synchronize(self.device)
        graph = ctypes.c_void_p()
        g0, i0, h0, s0 = gates.data_ptr(), its.data_ptr(), hist.data_ptr(), stamps.data_ptr()
        with torch.cuda.device(self.device), torch.cuda.stream(self.origin):
            _lib.call("tr_graph_begin", h)
            try:
                _lib.call("tr_clock_stamp", s0, h)  # stamps[0]: start of the replay
                for st in self.streams:
                    st.wait_stream(self.origin)
                for _ in range(chunk):
                    for i, s in enumerate(self.solvers):
                        st = self.streams[i]
                        with torch.cuda.stream(st):
                            _lib.call("tr_gate_set", g0 + 4 * i)
                            try:
                                s.pdhg_launch(0)
                                _lib.call("tr_stop_rule", s._red.data_ptr(), float(s.tol), int(s.max_iter),
                                          rules[i], g0 + 4 * i, i0 + 4 * i, h0 + 8 * maxit * i, s0 + 8 * (i + 1),
                                          st.cuda_stream)
                            finally:
                                _lib.call("tr_gate_set", None)
                for st in self.streams:
                    self.origin.wait_stream(st)
            finally:
                rc = lib.tr_graph_end(h, ctypes.byref(graph))
        _lib.check(rc, "tr_graph_end")
        self.captures += 1
        return graph

    def _run_persistent(self, chunk):
        import torch
        dev = self.device
        n = len(self.solvers)
        maxit = max(int(s.max_iter) for s in self.solvers)
        chunk = maxit if chunk is None else max(1, min(int(chunk), maxit))
        gates = torch.zeros(n, dtype=torch.int32, device=dev)
        its = torch.zeros(n, dtype=torch.int32, device=dev)
        hist = torch.zeros((n, maxit), dtype=torch.float64, device=dev)
        stamps = torch.zeros(n + 1, dtype=torch.int64, device=dev)
        rules = [1 if getattr(s, "algorithm", "") == "tvmw" else 0 for s in self.solvers]
        self._prepare()
        hz = ctypes.c_int64()
        with torch.cuda.device(dev):
            _lib.call("tr_clock_rate", ctypes.byref(hz))
        graph = self._capture_persistent(chunk, gates, its, hist, stamps, rules, maxit)
        lib = _lib.load()
        self.replays = 0
        t_first = None
        try:
            while True:
                _lib.call("tr_graph_launch", graph, self.origin.cuda_stream)
                self.replays += 1
                self.origin.synchronize()
                if t_first is None:
                    t_first = int(stamps[0].item())  # each replay restamps; the first one starts the clock
                if bool(torch.all(gates != 0)):
                    break
                if self.replays * chunk >= maxit:
                    raise RuntimeError("batched solve did not close every gate within the iteration caps")
        finally:
            lib.tr_graph_destroy(graph)
        its_h = its.tolist()
        hist_h = hist.cpu().numpy()
        st_h = stamps.tolist()
        out = []
        for i, s in enumerate(self.solvers):
            s.iterations = int(its_h[i])
            s.history = [float(v) for v in hist_h[i, :s.iterations]]
            s.elapsed = (st_h[i + 1] - t_first) / float(hz.value)
            out.append({"iterations": s.iterations, "seconds": s.elapsed, "history": s.history})
        return out

    def run(self, on_replay=None, mode="graph", chunk=10):
        """Advance every instance to its stopping rule.  Returns per-instance dicts with the
        iteration count, the seconds until it finished (from the first replay) and its history;
        each solver's ``ubar`` holds its result.  on_replay(k, active) runs after replay k (graph
        and streams modes); chunk is the iterations per replay of the persistent mode (None: the
        largest iteration cap, i.e. one launch)."""
        import torch
        if mode not in ("graph", "streams", "persistent"):
            raise ValueError(f"unknown mode {mode!r}")
        if mode != "streams" and any(getattr(s, "_bstreams", None) for s in self.solvers):
            # per-bin streams forked inside a capture are not supported: set_bin_streams(1) first
            raise ValueError(f"{mode} mode needs every instance on one bin stream (set_bin_streams(1))")
        if mode == "persistent":
            return self._run_persistent(chunk)
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
