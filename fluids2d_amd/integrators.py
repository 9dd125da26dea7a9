"""GPU Runge-Kutta integrator with the reference's interface
(integrators.py:9-17, 63-79): ``step(state, time)``, ``rhs``, ``diag``,
``scratch``.

The integrator owns a device twin of the state (``dstate``).  The public
methods keep the reference's semantics on the host numpy ``State``:

* ``diag(s)`` / ``rhs(s, ds)`` copy ``s`` to the device, run the model's
  operators there and copy the results back;
* ``step(state, time)`` copies the state in, advances one step on the device,
  copies it back and pushes the time forward.

``Model.run`` uses the device path directly (``begin`` / ``device_step`` /
``end``), copying to the host only at the points where host code reads the
state (history output, callbacks, diagnostics) and at the end.

Time step: the CFL step of ``model.py:84-87`` is computed on the device by the
kernel that produces the velocities at the end of each step; the host reads
the one float64 it needs for ``Time.pushforward``.
"""
import numpy as np
import torch

from . import _lib
from . import operators as op
from .equations import get_device_model
from .states import (Prognostic, allocate_device, download, get_specs, leaves,
                     upload)

rkdocs = {
    "rk3": "Strongly Stably Preserving RK3 (Shu 2003)",
    "ef": "Euler Forward",
    "enrk3": "Energy-Preserving RK3 (Celledoni et al. 2009)",
}


def get_integrator(param, mesh, state):
    """integrators.py:9-17"""
    if param.integrator in _lib.NSTAGES:
        return RKIntegrator(param, mesh, state)
    if param.integrator == "LFRA":
        return LFRAIntegrator(param, mesh, state)
    raise NotImplementedError(f"{param.integrator} is not implemented on the GPU path")


class RKIntegrator:
    """Runge-Kutta integrator"""

    def __init__(self, param, mesh, state):
        self.param = param
        self.mesh = mesh
        self.nstages = _lib.NSTAGES[param.integrator]
        self.scratch = [Prognostic(param, mesh.shape) for _ in range(self.nstages)]
        self.dtype = torch.float64 if param.precision == "fp64" else torch.float32
        dev = mesh.device
        specs = get_specs(param)
        self.specs = specs
        self.dstate = allocate_device(specs.variables, mesh.shape, self.dtype, dev)
        self.dalt = allocate_device(specs.prognostic, mesh.shape, self.dtype, dev)
        self.dscratch = [allocate_device(specs.prognostic, mesh.shape, self.dtype, dev)
                         for _ in range(self.nstages)]
        self.model = get_device_model(param, mesh)
        self.d_dt = torch.zeros(1, dtype=torch.float64, device=dev)
        self.d_dtn = torch.zeros(1, dtype=torch.float64, device=dev)
        self._dt_written = None  # (value, buffer) of the last fixed dt written
        self._h_dt = torch.zeros(1, dtype=torch.float64).pin_memory()
        self._dtn_valid = False
        # Models with ``defers_cfl``: the last stage leaves the next step's
        # CFL step as block maxima in _cfl (_dtn_part: d_dtn is not written
        # yet); after the swap (_dt_part) the first stage of the step forms
        # it from them and writes d_dt.  Any other reader of the dt buffers
        # finishes it first (_finish_cfl).
        self._cfl = None
        self._dtn_part = False
        self._dt_part = False
        self.__doc__ = (self.__doc__ or "") + f"\n{rkdocs[param.integrator]}"
        self.rhs = self._host_rhs
        self.diag = self._host_diag
        self._graphs = {}
        self._graph_solves = []
        self._glog = None
        self._eager_steps = 0
        self._graph_mesh = None
        self.forcings = []

    # -- forcing (model.py:121-123, equations.py:229-238) -------------------------
    def add_forcing(self, forcing):
        """forcing(param, mesh, s, ds) is called after every rhs, with the
        device state and tendency: torch tensors of the padded layout, on
        which slicing, scalar arithmetic and numpy-array operands work as on
        the reference's numpy arrays (a numpy operand is copied to the
        device at each call; a forcing that keeps its arrays as device
        tensors avoids the copy).  Stages then run unfused (rhs, forcing,
        addto, diag) and steps are not graph-captured (the forcing is host
        code).  A forcing that reads ``mesh.time`` sees the time at the
        start of the step, as in the reference, when the steps are taken
        by ``Model.run`` / ``device_step``."""
        self.forcings.append(forcing)

    def _force(self, s, ds):
        """The forcing functions, on views of the device state and tendency
        that accept numpy operands (see _HostOperands)."""
        if not self.forcings:
            return
        sv, dsv = _host_operands(s), _host_operands(ds)
        for f in self.forcings:
            f(self.param, self.mesh, sv, dsv)

    def _forced_stage(self, s, k, dtn):
        """One RK stage in the reference's order (integrators.py:92-107):
        rhs, forcing terms, addto over the prognostic arrays, diag."""
        p = self.param
        ds = self.dscratch[k]
        self.model.rhs(s, ds)
        self._force(s, ds)
        ys = self._prog(s)
        prev = [leaves(self.dscratch[j]) for j in range(k + 1)]
        for n, y in enumerate(ys):
            op.rk_addto(y, [pv[n] for pv in prev], p.integrator, k, self.d_dt)
        self.model.diag(s)
        if dtn is not None:
            op.cfl_dt(self.model.cfl_velocity(s), p.cfl, p.dtmax, dtn)

    # -- host API (reference semantics) ----------------------------------------
    def _prog(self, s):
        """The prognostic arrays of state s, in field order."""
        return [a for n in self.specs.prognostic for a in leaves(getattr(s, n))]

    def _host_rhs(self, s, ds):
        upload(s, self.dstate)
        dds = self.dscratch[0]
        self.model.rhs(self.dstate, dds)
        download(dds, ds)
        if "flx" in self.dstate._fields and hasattr(s, "flx"):
            download(self.dstate, s, names=["flx"])

    def _host_diag(self, s):
        upload(s, self.dstate)
        self.model.diag(self.dstate)
        self.model.materialize(self.dstate)
        download(self.dstate, s)
        self._dtn_valid = self._dtn_part = self._dt_part = False

    def step(self, state, time):
        """RKIntegrator.step (integrators.py:76-79): one step of time.dt on
        the host state (copy in, advance, copy out).  As in the reference,
        dt is the time's current one; Model.set_dt recomputes it."""
        self.begin(state)
        self._write_dt(time.dt)
        self.device_step(time)
        self.end(state)

    # -- device path --------------------------------------------------------------
    def begin(self, state):
        """Host state -> device twin."""
        upload(state, self.dstate)
        self._dtn_valid = self._dtn_part = self._dt_part = False

    def end(self, state, names=None):
        """Device twin -> host state."""
        self.model.materialize(self.dstate, names)
        download(self.dstate, state, names)

    def set_dt(self, time):
        """Model.set_dt (model.py:71-87): dt on the device and on the host."""
        p = self.param
        if p.dt > 0:
            time.dt = p.dt
            self._write_dt(p.dt)
            return
        if p.model == "rsw":
            c = (p.g * p.H) ** 0.5
            maxU = c / self.mesh.dx + c / self.mesh.dy
            time.dt = min(p.cfl / maxU, p.dtmax)
            self._write_dt(time.dt)
            return
        self._dt_written = None
        if not self._dtn_valid:
            op.cfl_dt(self.model.cfl_velocity(self.dstate), p.cfl, p.dtmax, self.d_dtn)
            self._dtn_valid = True
        elif self._dtn_part:
            self._finish_cfl(self.d_dtn)
        self._dtn_part = self._dt_part = False
        self.d_dt, self.d_dtn = self.d_dtn, self.d_dt
        self._h_dt.copy_(self.d_dt, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        time.dt = float(self._h_dt[0])

    def _write_dt(self, value):
        """A fixed dt into d_dt.  Skipped when d_dt already holds it (it was
        the last value written to this buffer and the CFL path, which swaps
        and overwrites the dt buffers, has not run since): a step graph of a
        fixed-dt run then carries no set kernel."""
        key = (float(value), self.d_dt.data_ptr())
        if self._dt_written != key:
            op.set_scalar(self.d_dt, value)
            self._dt_written = key

    def _wants_dtn(self):
        p = self.param
        return not (p.dt > 0 or p.model == "rsw")

    def _defers(self):
        return self._wants_dtn() and getattr(self.model, "defers_cfl", False) \
            and not self.forcings

    def _finish_cfl(self, dt_dev):
        """The CFL step pending in _cfl into dt_dev (one small launch)."""
        p = self.param
        op.cfl_finish(self._cfl, dt_dev, p.cfl, p.dtmax)

    def _dt_ready(self):
        """d_dt written, for a path that reads it directly."""
        if self._dt_part:
            self._finish_cfl(self.d_dt)
            self._dt_part = False

    def nonfinite(self):
        """Names of the prognostic fields of the device state with a
        non-finite value on a wet cell (one flag per field, read on the host:
        a synchronisation point)."""
        names = [n for n in self.specs.prognostic]
        flags = torch.zeros(len(names), dtype=torch.int32, device=self.d_dt.device)
        m = self.mesh
        for k, n in enumerate(names):
            for a in leaves(getattr(self.dstate, n)):
                _lib.call("f2d_nonfinite", _lib.dtype_code(a.dtype), m.grid_ref(), _lib.ptr(a),
                          _lib.ptr(m.d_msk), _lib.ptr(flags[k:k + 1]), _lib.stream_ptr())
        return [n for n, f in zip(names, flags.cpu().tolist()) if f]

    def _log_dt(self, dt_log):
        """The step's dt into dt_log by a kernel of its own (models whose
        first stage does not log it)."""
        _lib.call("f2d_copy_f64", _lib.ptr(dt_log), _lib.ptr(self.d_dt), 1, _lib.stream_ptr())

    def advance(self, stats=True, dt_log=None):
        """One RK step on the device with the dt already in d_dt; dt_log (a
        one-element float64 device tensor) receives that dt."""
        s, alt = self.dstate, self.dalt
        dtn = self.d_dtn if self._wants_dtn() else None
        if self.forcings:
            self._dt_ready()
            if dt_log is not None:
                self._log_dt(dt_log)
            for k in range(self.nstages):
                self._forced_stage(s, k, dtn if k == self.nstages - 1 else None)
            self._dtn_valid = dtn is not None
            self._dtn_part = False
            return
        defer = self._defers()
        if not defer:
            self._dt_ready()
        elif self._cfl is None:
            if torch.cuda.is_current_stream_capturing():
                raise RuntimeError("the CFL partials must be allocated before a capture")
            self._cfl = torch.zeros(op.cfl_partials(self.mesh), dtype=torch.float64,
                                    device=self.d_dt.device)
        logs = getattr(self.model, "logs_dt", False)
        if dt_log is not None and not logs:
            self._log_dt(dt_log)
        for k in range(self.nstages):
            ds_out = self.dscratch[k] if k < self.nstages - 1 else None
            prev = [self.dscratch[j] for j in range(k)]
            last = k == self.nstages - 1
            kw = {"dt_log": dt_log} if (k == 0 and logs and dt_log is not None) else {}
            if k == 0 and self._dt_part:
                kw["cfl_in"] = self._cfl
            if last and defer:
                kw["cfl_out"] = self._cfl
            s, alt = self.model.stage(s, alt, [_one(x) for x in prev],
                                      _one(ds_out) if ds_out is not None else None,
                                      k, self.d_dt, None if (defer or not last) else dtn,
                                      stats, **kw)
        self.dstate, self.dalt = s, alt
        self._dt_part = False
        self._dtn_valid = dtn is not None
        self._dtn_part = defer

    def staged(self, dt_log=None):
        """advance(stats=False, dt_log=dt_log) cut at each stage's elliptic
        solve, for a caller that batches the solves of many models (the
        model must offer ``stage_split``; no forcing): per stage, the
        (pre, post, cfl) callables of the model (cfl None but on the last
        stage).  Running pre, the solve, post, cfl for k = 0, 1, ... is the
        step.  The dt is the one in d_dt now (call after _select_dt)."""
        assert not self.forcings and hasattr(self.model, "stage_split")
        self._dt_ready()
        if dt_log is not None and not getattr(self.model, "logs_dt", False):
            self._log_dt(dt_log)
            dt_log = None
        s = self.dstate
        dtn = self.d_dtn if self._wants_dtn() else None
        plan = []
        for k in range(self.nstages):
            ds_out = _one(self.dscratch[k]) if k < self.nstages - 1 else None
            prev = [_one(self.dscratch[j]) for j in range(k)]
            last = k == self.nstages - 1
            plan.append(self.model.stage_split(s, prev, ds_out, k, self.d_dt,
                                               dtn if last else None,
                                               dt_log if k == 0 else None))
        self._dtn_valid = dtn is not None
        self._dtn_part = False
        return plan

    def device_step(self, time):
        self.advance()
        time.pushforward()

    def _select_dt(self, defer=False):
        """The dt of the next step into d_dt, on the device (Model.set_dt,
        model.py:71-87).  defer (for a following ``advance``): a CFL step
        pending as block maxima stays so, and the step's first stage forms
        it."""
        p = self.param
        self._dt_part = False
        if p.dt > 0:
            self._write_dt(p.dt)
        elif p.model == "rsw":
            c = (p.g * p.H) ** 0.5
            self._write_dt(min(p.cfl / (c / self.mesh.dx + c / self.mesh.dy), p.dtmax))
        else:
            self._dt_written = None
            if not self._dtn_valid:
                op.cfl_dt(self.model.cfl_velocity(self.dstate), p.cfl, p.dtmax, self.d_dtn)
                self._dtn_part = False
            elif self._dtn_part and not (defer and self._defers()):
                self._finish_cfl(self.d_dtn)
                self._dtn_part = False
            self.d_dt, self.d_dtn = self.d_dtn, self.d_dt
            self._dt_part, self._dtn_part = self._dtn_part, False

    def device_steps(self, nsteps, stats=False, graph=None):
        """nsteps steps with dt kept on the device (no host round trip).
        Returns the per-step dt as a device tensor.

        graph (default ``param.hipgraph``): after two eager steps, runs of
        ``graph_steps`` steps are captured once into a HIP graph and
        replayed; the launches are the eager ones, so results are
        identical."""
        dts = torch.zeros(nsteps, dtype=torch.float64, device=self.d_dt.device)
        use = getattr(self.param, "hipgraph", True) if graph is None else graph
        use = use and not stats and getattr(self.model, "capturable", True) \
            and not self.forcings
        G = self.graph_steps
        n = 0
        while n < nsteps:
            if use and nsteps - n >= G and self._eager_steps >= 2 and \
                    (self._dtn_valid or not self._wants_dtn()) and \
                    (self._dtn_part or not self._defers()):
                g = self._graph()
                g.replay()
                for e, k in self._graph_solves:
                    e.graph_replayed(k, 0)
                _lib.call("f2d_copy_f64", _lib.ptr(dts[n:n + G]), _lib.ptr(self._glog), G,
                          _lib.stream_ptr())
                n += G
                continue
            self._select_dt(defer=True)
            self.advance(stats=stats, dt_log=dts[n:n + 1])
            self._eager_steps += 1
            n += 1
        return dts

    def prepare_graph(self):
        """Run the eager steps a capture needs and capture the step graph of
        the current buffer assignment now, so that a following
        ``device_steps`` replays only (nothing is captured inside it when its
        step count is a multiple of ``graph_steps``).  Returns the number of
        steps taken (0 to 2 eager steps; the capture itself advances
        nothing)."""
        n = 0
        while self._eager_steps < 2 or (self._wants_dtn() and not self._dtn_valid) or \
                (self._defers() and not self._dtn_part):
            self._select_dt(defer=True)
            self.advance(stats=False)
            self._eager_steps += 1
            n += 1
        self._graph()
        return n

    def graph_count(self):
        """Number of step graphs captured so far (for callers that check that
        a timed region captured none)."""
        return len(self._graphs)

    # -- HIP graphs -----------------------------------------------------------------
    graph_steps = 4  # steps per captured graph; even, so ping-pong buffers return

    def _graph_key(self):
        if self._graph_mesh != self.mesh.version:
            # a re-finalised mesh has new solvers and masks: drop old graphs
            self._graphs = {}
            self._graph_mesh = self.mesh.version
        return (self.mesh.version, id(self.model),
                tuple(t.data_ptr() for t in leaves(self.dstate)),
                tuple(t.data_ptr() for t in leaves(self.dalt)),
                self.d_dt.data_ptr(), self.d_dtn.data_ptr(), self._dtn_part)

    def _solvers(self):
        m = self.mesh
        out = []
        for name in ("poisson_vertices", "poisson_centers", "qg_helmholtz"):
            e = m.__dict__.get(name)
            if e is not None:
                out.append(e)
        return out

    def _graph(self):
        """The captured graph of graph_steps steps for the current buffer
        assignment (captured on first use)."""
        key = self._graph_key()
        g = self._graphs.get(key)
        if g is not None:
            return g
        G = self.graph_steps
        if self._glog is None:
            self._glog = torch.zeros(G, dtype=torch.float64, device=self.d_dt.device)
        sols = self._solvers()
        before = [e.calls for e in sols]
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with _lib.capture(g):
            for i in range(G):
                self._select_dt(defer=True)
                self.advance(stats=False, dt_log=self._glog[i:i + 1])
        if self._graph_key() != key:
            raise RuntimeError("graph capture did not return the state buffers to "
                               "their roles (odd number of buffer swaps)")
        self._graph_solves = [(e, e.calls - c0) for e, c0 in zip(sols, before)]
        for e, k in self._graph_solves:
            e.graph_replayed(0, k)
        self._graphs[key] = g
        return g

class LFRAIntegrator(RKIntegrator):
    """Leap-Frog integrator with a Robert-Asselin filter (integrators.py:20-53):
    the first iteration (ite == 0) is Euler forward, then
      sa += 2 dt ds;  s += gamma sa + gamma sb - 2 gamma s;
      (sa, s, sb) <- (s, sa, s)   (rightpermute, integrators.py:134-137)
    with sa, sb the filtered previous state and ds = rhs(s).  The state's
    prognostic arrays, the two levels and the tendency stay on the device;
    the addto is f2d_addto (dt from the device)."""

    def __init__(self, param, mesh, state):
        param_integ = param.integrator
        param.integrator = "ef"  # RKIntegrator's buffers: one stage
        try:
            super().__init__(param, mesh, state)
        finally:
            param.integrator = param_integ
        self.gamma = float(param.RAgamma)
        dev = mesh.device
        specs = self.specs
        self._sa = allocate_device(specs.prognostic, mesh.shape, self.dtype, dev)
        self._sb = allocate_device(specs.prognostic, mesh.shape, self.dtype, dev)
        self._first = True
        self.__doc__ = LFRAIntegrator.__doc__

    def _addto(self, y, terms, dtmask):
        cs = [c for c, _ in terms] + [0.0] * (3 - len(terms))
        xs = [x for _, x in terms] + [None] * (3 - len(terms))
        _lib.call("f2d_addto", _lib.dtype_code(y.dtype), _lib.ptr(y), _lib.ptr(xs[0]),
                  _lib.ptr(xs[1]), _lib.ptr(xs[2]), y.numel(), len(terms), float(cs[0]),
                  float(cs[1]), float(cs[2]), _lib.ptr(self.d_dt), dtmask, _lib.stream_ptr())

    def _defers(self):
        return False  # the step's dt comes from f2d_cfl_dt

    def advance(self, stats=True, dt_log=None):
        p = self.param
        s = self.dstate
        dtn = self.d_dtn if self._wants_dtn() else None
        self._dt_ready()
        if dt_log is not None:
            self._log_dt(dt_log)
        ds = self.dscratch[0]
        self.model.rhs(s, ds)
        self._force(s, ds)
        ys, dsl = self._prog(s), leaves(ds)
        sa, sb = leaves(self._sa), leaves(self._sb)
        if self._first:
            for y, a, b, d in zip(ys, sa, sb, dsl):
                b.copy_(y)
                a.copy_(y)
                self._addto(y, [(1.0, d)], 1)          # addto(s, dt, ds)
            self._first = False
        else:
            g = self.gamma
            for y, a, b, d in zip(ys, sa, sb, dsl):
                self._addto(a, [(2.0, d)], 1)          # addto(sa, 2*dt, ds)
                self._addto(y, [(g, a), (g, b), (-2 * g, y)], 0)
                b.copy_(y)                             # rightpermute(sa, s, sb)
                y.copy_(a)
                a.copy_(b)
        self.model.diag(s)
        if dtn is not None:
            op.cfl_dt(self.model.cfl_velocity(s), p.cfl, p.dtmax, dtn)
        self._dtn_valid = dtn is not None
        self._dtn_part = False

    def device_step(self, time):
        self._first = time.ite == 0
        self.advance()
        time.pushforward()

    def device_steps(self, nsteps, stats=False, graph=None):
        return super().device_steps(nsteps, stats=stats, graph=False)


class _HostOperands(torch.Tensor):
    """A device tensor whose arithmetic accepts numpy arrays as operands
    (copied to the tensor's device first): forcing functions written for
    the reference's numpy states, e.g. ``ds.pv[:, :] += curltau`` with
    ``curltau`` a numpy array, run unchanged on the device state.  numpy
    defers to these operators (``__array_ufunc__ = None``), so an array on
    the left (``curl * s.q``) is converted as well."""

    __array_ufunc__ = None

    def _operand(self, other):
        if isinstance(other, np.ndarray):
            with torch._C.DisableTorchFunctionSubclass():
                dev = self.device
            return torch.as_tensor(other, device=dev)
        return other

    def __setitem__(self, index, value):
        return super().__setitem__(index, self._operand(value))


def _with_operand(name):
    base = getattr(torch.Tensor, name)

    def op(self, other):
        return base(self, self._operand(other))
    op.__name__ = name
    return op


for _name in ("add", "sub", "mul", "truediv", "pow"):
    for _form in ("__%s__", "__r%s__", "__i%s__"):
        _n = _form % _name
        if hasattr(torch.Tensor, _n):
            setattr(_HostOperands, _n, _with_operand(_n))


def _host_operands(x):
    if isinstance(x, torch.Tensor):
        return x.as_subclass(_HostOperands)
    return type(x)(*[_host_operands(v) for v in x])


def _one(x):
    """The single prognostic array of a one-field prognostic namedtuple."""
    if x is None:
        return None
    return x[0] if len(x) == 1 else x
