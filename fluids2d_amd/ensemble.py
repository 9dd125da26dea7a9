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

Member batching (``batch=True``, the default): members whose stages pause
at a Helmholtz solve of the capacitance method in FP32 (qgrsw) are grouped
by operator (same grid, spacings, f0/g/H, periodicity, methods) and each
group is stepped stage by stage as one: per stage, one launch per kernel
for the whole group before the solve (``LaunchBatch``: the members' entry
point calls recorded and merged, blockIdx.y = member), the group's solve
as one batch (``elliptic.PoissonBatch``), one launch per kernel after it;
then each member's CFL reduction.  Groups run side by side on the streams,
members outside any group take whole steps on their lanes.  Every member's
kernels keep their arithmetic and order, so the results are still those of
its solo run, bit for bit.
"""
import ctypes
import time

import torch

from . import _lib
from .elliptic import PoissonBatch, _flush_pending, _release


class LaunchBatch:
    """The kernel launches of ``parts`` (one callable per member, each making
    the same entry point calls on its own arrays), recorded and merged into
    one launch per kernel for all members (f2d_record_begin / end,
    f2d_launch_batch_create / run)."""

    def __init__(self, parts):
        recs = []
        self._h = None
        try:
            for fn in parts:
                _lib.call("f2d_record_begin")
                h = ctypes.c_void_p()
                try:
                    fn()
                finally:
                    _lib.call("f2d_record_end", ctypes.byref(h))
                    recs.append(h)
            self.launches = _lib.lib().f2d_recording_size(recs[0])
            arr = ctypes.c_void_p * len(recs)
            h = ctypes.c_void_p()
            _flush_pending()
            _lib.call("f2d_launch_batch_create", arr(*recs), len(recs), ctypes.byref(h))
            self._h = h
        finally:
            for r in recs:
                _lib.lib().f2d_recording_destroy(r)

    def run(self):
        _lib.call("f2d_launch_batch_run", self._h, _lib.stream_ptr())

    def __del__(self):
        h = getattr(self, "_h", None)
        self._h = None
        _release(h, "f2d_launch_batch_destroy")


class Ensemble:
    def __init__(self, models, nstreams=8, graph_steps=2, hipgraph=True, batch=True):
        if graph_steps % 2:
            raise ValueError("graph_steps must be even (ping-pong buffers)")
        self.models = list(models)
        if not self.models:
            raise ValueError("empty ensemble")
        self.G = int(graph_steps)
        self.hipgraph = hipgraph
        self.batch = batch
        self._groups = None
        self._groups_key = None
        self._inb = set()
        self._lb = {}
        dev = self.models[0].mesh.device
        self.dtlog = torch.zeros((len(self.models), self.G), dtype=torch.float64, device=dev)
        self._slot = torch.zeros(len(self.models), dtype=torch.float64, device=dev)
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

    # -- member batching ---------------------------------------------------------------
    def _batchable(self, i):
        m = self.models[i]
        integ = m.integrator
        model = integ.model
        return (self.batch and m.param.integrator in _lib.NSTAGES and hasattr(integ, "staged")
                and not integ.forcings and hasattr(model, "stage_split")
                and getattr(model, "fused", False)
                and PoissonBatch.eligible(m.mesh.qg_helmholtz))

    def _group_key(self, i):
        m = self.models[i]
        p = m.param
        return (type(m.integrator.model).__name__, PoissonBatch.key(m.mesh.qg_helmholtz),
                p.integrator, p.compflux, p.vortexforce, p.innerproduct, p.f0, p.g, p.H,
                m.mesh.grid.tile, getattr(p, "poisson_refine", "every"))

    def groups(self):
        """The member groups: a list of (member indices, PoissonBatch), one
        per operator, the solves bound to the members' current pv and psi
        (rebuilt, with the groups' launch batches, when a member's solver
        or arrays change)."""
        idx = [i for i in range(len(self.models)) if self._batchable(i)]
        key = tuple((i, id(self.models[i].mesh.qg_helmholtz), self._group_key(i),
                     self.models[i].integrator.dstate.pv.data_ptr(),
                     self.models[i].integrator.dstate.psi.data_ptr()) for i in idx)
        if self._groups is None or self._groups_key != key:
            by = {}
            for i in idx:
                by.setdefault(self._group_key(i), []).append(i)
            groups = []
            for members in by.values():
                solvers = [self.models[i].mesh.qg_helmholtz for i in members]
                st = [self.models[i].integrator.dstate for i in members]
                groups.append((members, PoissonBatch(solvers, [x.pv for x in st],
                                                     [x.psi for x in st])))
            self._groups, self._groups_key = groups, key
            self._inb = {i for members, _ in groups for i in members}
            self._lb = {}
        return self._groups

    def _batched_step(self, col):
        """One step of every member, the groups stage by stage with one
        launch per kernel; member i's dt into col[i]."""
        groups = self.groups()
        bodies = [[] for _ in self.streams]
        for g in range(len(groups)):
            bodies[g % len(bodies)].append(lambda g=g: self._group_step(g))
        for j, lane in enumerate(self.lanes):
            for i in lane:
                if i not in self._inb:
                    bodies[j].append(lambda i=i: self._member_step(i, self._slot[i:i + 1]))
        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        for body, s in zip(bodies, self.streams):
            with torch.cuda.stream(s):
                for b in body:
                    b()
        for s in self.streams:
            main.wait_stream(s)
        col.copy_(self._slot)

    def _group_step(self, g):
        members, solve = self._groups[g]
        plans = []
        for i in members:
            integ = self.models[i].integrator
            integ._select_dt()
            plans.append(integ.staged(dt_log=self._slot[i:i + 1]))
        nst = len(plans[0])
        dts = tuple(self.models[i].integrator.d_dt.data_ptr() for i in members)

        def part(plan, k):
            if k > 0:
                plan[k - 1][1]()  # after stage k-1's solve
            if k < nst:
                plan[k][0]()      # before stage k's solve

        model = self.models[members[0]].integrator.model
        for k in range(nst + 1):
            self._launch((g, k, dts), [lambda p=p, k=k: part(p, k) for p in plans])
            if k < nst:
                solve.solve(nref=model.refine_for(k))
        cfls = [p[-1][2] for p in plans if p[-1][2] is not None]
        if cfls:  # the members' CFL steps, one launch per kernel
            self._launch((g, "cfl", dts), cfls)

    def _launch(self, key, parts):
        """The members' parts as one launch per kernel: the batch recorded
        for this key (recorded now, outside a capture), or the parts one by
        one where they cannot be recorded."""
        if key not in self._lb:
            if _lib.capturing():
                raise RuntimeError("ensemble: a member batch was not recorded before the capture")
            try:
                self._lb[key] = LaunchBatch(parts)
            except _lib.F2DError:
                self._lb[key] = None  # an entry point that cannot be recorded
        lb = self._lb[key]
        if lb is None:
            for fn in parts:
                fn()
        else:
            lb.run()

    # -- stepping -------------------------------------------------------------------
    def _member_step(self, i, dst):
        integ = self.models[i].integrator
        integ._select_dt(defer=True)
        integ.advance(stats=False, dt_log=dst)

    def _step_all(self, col):
        """One step of every member; member i's dt into col[i]."""
        if self.batch and self.groups():
            self._batched_step(col)
        else:
            self._fork(lambda i: self._member_step(i, col[i:i + 1]))

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
        if self.batch:
            self.groups()  # set up (device tables) before the capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with _lib.capture(g):
            if self.batch and self.groups():
                for n in range(self.G):
                    self._step_all(self.dtlog[:, n])
            else:
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
        # (a member whose next CFL step is pending as block maxima captures
        # its steps in that state; see RKIntegrator._select_dt)
        return self._eager >= 2 and all(
            (m.integrator._dtn_valid or not m.integrator._wants_dtn()) and
            (m.integrator._dtn_part or not m.integrator._defers()) for m in self.models)

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
            self._step_all(dts[:, n])
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
