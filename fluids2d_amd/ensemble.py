"""Ensembles: many independent models, of uneven grid sizes and masks,
advanced together on one GPU.

Each member is a ``Model`` with its own mesh, device state, solvers and
CFL time step (the reference runs one ``Model`` per process; an ensemble is
that loop over members, model.py:29-69).  Members are bin-packed by cells
onto S streams (longest first onto the least loaded stream), so that the
small launches of small members and the coarse multigrid levels of large
ones run side by side on the 256 CUs.  After two eager steps, G steps of
all members are captured once into one HIP graph -- a fork into the S
streams, each stream's members one after another, a join -- and that one
graph is replayed: one launch advances the whole ensemble by G steps with
no host work in between.

Results do not depend on S, on the packing or on the graph: each member's
kernels and their order are those of its own ``device_steps``.  Reductions
of the CFL step use per-stream scratch (``f2d_reserve_stream_scratch``),
so concurrent members share no buffer.
"""
import time

import torch

from . import _lib


class Ensemble:
    def __init__(self, models, nstreams=8, graph_steps=2, hipgraph=True):
        if graph_steps % 2:
            raise ValueError("graph_steps must be even (ping-pong buffers)")
        self.models = list(models)
        if not self.models:
            raise ValueError("empty ensemble")
        self.G = int(graph_steps)
        self.hipgraph = hipgraph
        dev = self.models[0].mesh.device
        self.dtlog = torch.zeros((len(self.models), self.G), dtype=torch.float64, device=dev)
        self._eager = 0
        self._set_streams(nstreams)

    # -- layout -------------------------------------------------------------------
    def cells(self, i):
        m = self.models[i].mesh
        return m.nx * m.ny

    def _set_streams(self, nstreams):
        S = max(1, min(int(nstreams), len(self.models)))
        self.nstreams = S
        self.streams = [torch.cuda.Stream() for _ in range(S)]
        for s in self.streams:
            _lib.call("f2d_reserve_stream_scratch", _lib.stream_ptr(s))
        load = [0] * S
        self.lanes = [[] for _ in range(S)]
        for i in sorted(range(len(self.models)), key=self.cells, reverse=True):
            k = min(range(S), key=lambda j: load[j])
            self.lanes[k].append(i)
            load[k] += self.cells(i)
        self.graph = None
        self._key = None

    # -- stepping -------------------------------------------------------------------
    def _member_step(self, i, dst):
        integ = self.models[i].integrator
        integ._select_dt()
        integ.advance(stats=False, dt_log=dst)

    def _fork(self, body):
        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        for lane, s in zip(self.lanes, self.streams):
            with torch.cuda.stream(s):
                for i in lane:
                    body(i)
        for s in self.streams:
            main.wait_stream(s)

    def _graph_key(self):
        return tuple(m.integrator._graph_key() for m in self.models)

    def _capture(self):
        key = self._graph_key()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with _lib.capture(g):
            def body(i):
                for n in range(self.G):
                    self._member_step(i, self.dtlog[i, n:n + 1])
            self._fork(body)
        if self._graph_key() != key:
            raise RuntimeError("ensemble capture did not return the buffers to their roles")
        self.graph, self._key = g, key

    def _ready(self):
        if not all(getattr(m.integrator.model, "capturable", True) for m in self.models):
            return False  # a member's steps check on the host: eager only
        if any(getattr(m.integrator, "forcings", None) for m in self.models):
            return False  # host forcing code runs at every step: eager only
        return self._eager >= 2 and all(
            m.integrator._dtn_valid or not m.integrator._wants_dtn() for m in self.models)

    def steps(self, nsteps):
        """Advance every member by nsteps steps; returns the (members, nsteps)
        device tensor of their time steps."""
        dts = torch.zeros((len(self.models), nsteps), dtype=torch.float64,
                          device=self.dtlog.device)
        n = 0
        while n < nsteps:
            if self.hipgraph and nsteps - n >= self.G and self._ready():
                if self.graph is None or self._key != self._graph_key():
                    self._capture()
                self.graph.replay()
                dts[:, n:n + self.G].copy_(self.dtlog)
                n += self.G
                continue
            col = n
            self._fork(lambda i: self._member_step(i, dts[i, col:col + 1]))
            self._eager += 1
            n += 1
        return dts

    def run(self, nsteps):
        """steps(), then each member's Time pushed forward by its dts."""
        dts = self.steps(nsteps).cpu().numpy()
        for m, d in zip(self.models, dts):
            for v in d:
                m.time.dt = float(v)
                m.time.pushforward()
        return dts

    def end(self):
        """Device twins -> host states."""
        for m in self.models:
            m.integrator.end(m.state)

    # -- tuning -------------------------------------------------------------------
    def tune_streams(self, candidates=(2, 4, 8, 16), reps=2):
        """Pick the stream count by timing replays of the ensemble graph
        (each replay advances the members by G steps: the caller's run
        continues from there).  Returns {S: ms per step}."""
        if not self._ready():
            self.steps(2)
        times = {}
        for S in candidates:
            if S > len(self.models):
                continue
            self._set_streams(S)
            self._capture()
            self.graph.replay()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                self.graph.replay()
            torch.cuda.synchronize()
            times[S] = (time.perf_counter() - t0) * 1e3 / (reps * self.G)
        best = min(times, key=times.get)
        self._set_streams(best)
        return times
