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
        self._h_dt = torch.zeros(1, dtype=torch.float64).pin_memory()
        self._dtn_valid = False
        self.__doc__ = (self.__doc__ or "") + f"\n{rkdocs[param.integrator]}"
        self.rhs = self._host_rhs
        self.diag = self._host_diag
        self._graphs = {}
        self._graph_solves = []
        self._glog = None
        self._eager_steps = 0
        self._graph_mesh = None

    # -- host API (reference semantics) ----------------------------------------
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
        self._dtn_valid = False

    def step(self, state, time):
        """One step on the host state: copy in, advance, copy out."""
        self.begin(state)
        self.set_dt(time)
        self.device_step(time)
        self.end(state)

    # -- device path --------------------------------------------------------------
    def begin(self, state):
        """Host state -> device twin."""
        upload(state, self.dstate)
        self._dtn_valid = False

    def end(self, state, names=None):
        """Device twin -> host state."""
        self.model.materialize(self.dstate)
        download(self.dstate, state, names)

    def set_dt(self, time):
        """Model.set_dt (model.py:71-87): dt on the device and on the host."""
        p = self.param
        if p.dt > 0:
            time.dt = p.dt
            op.set_scalar(self.d_dt, p.dt)
            return
        if p.model == "rsw":
            c = (p.g * p.H) ** 0.5
            maxU = c / self.mesh.dx + c / self.mesh.dy
            time.dt = min(p.cfl / maxU, p.dtmax)
            op.set_scalar(self.d_dt, time.dt)
            return
        if not self._dtn_valid:
            op.cfl_dt(self.model.cfl_velocity(self.dstate), p.cfl, p.dtmax, self.d_dtn)
            self._dtn_valid = True
        self.d_dt, self.d_dtn = self.d_dtn, self.d_dt
        self._h_dt.copy_(self.d_dt, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        time.dt = float(self._h_dt[0])

    def _wants_dtn(self):
        p = self.param
        return not (p.dt > 0 or p.model == "rsw")

    def advance(self, stats=True):
        """One RK step on the device with the dt already in d_dt."""
        s, alt = self.dstate, self.dalt
        dtn = self.d_dtn if self._wants_dtn() else None
        for k in range(self.nstages):
            ds_out = self.dscratch[k] if k < self.nstages - 1 else None
            prev = [self.dscratch[j] for j in range(k)]
            last = k == self.nstages - 1
            s, alt = self.model.stage(s, alt, [_one(x) for x in prev],
                                      _one(ds_out) if ds_out is not None else None,
                                      k, self.d_dt, dtn if last else None, stats)
        self.dstate, self.dalt = s, alt
        self._dtn_valid = dtn is not None

    def device_step(self, time):
        self.advance()
        time.pushforward()

    def _select_dt(self):
        """The dt of the next step into d_dt, on the device (Model.set_dt,
        model.py:71-87)."""
        p = self.param
        if p.dt > 0:
            op.set_scalar(self.d_dt, p.dt)
        elif p.model == "rsw":
            c = (p.g * p.H) ** 0.5
            op.set_scalar(self.d_dt, min(p.cfl / (c / self.mesh.dx + c / self.mesh.dy),
                                         p.dtmax))
        else:
            if not self._dtn_valid:
                op.cfl_dt(self.model.cfl_velocity(self.dstate), p.cfl, p.dtmax, self.d_dtn)
            self.d_dt, self.d_dtn = self.d_dtn, self.d_dt

    def device_steps(self, nsteps, stats=False, graph=None):
        """nsteps steps with dt kept on the device (no host round trip).
        Returns the per-step dt as a device tensor.

        graph (default ``param.hipgraph``): after two eager steps, runs of
        ``graph_steps`` steps are captured once into a HIP graph and
        replayed; the launches are the eager ones, so results are
        identical."""
        dts = torch.zeros(nsteps, dtype=torch.float64, device=self.d_dt.device)
        use = getattr(self.param, "hipgraph", True) if graph is None else graph
        use = use and not stats and getattr(self.model, "capturable", True)
        G = self.graph_steps
        n = 0
        while n < nsteps:
            if use and nsteps - n >= G and self._eager_steps >= 2 and \
                    (self._dtn_valid or not self._wants_dtn()):
                g = self._graph()
                g.replay()
                for e, k in self._graph_solves:
                    e.graph_replayed(k, 0)
                _lib.call("f2d_copy_f64", _lib.ptr(dts[n:n + G]), _lib.ptr(self._glog), G,
                          _lib.stream_ptr())
                n += G
                continue
            self._select_dt()
            dts[n:n + 1].copy_(self.d_dt)
            self.advance(stats=stats)
            self._eager_steps += 1
            n += 1
        return dts

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
                self.d_dt.data_ptr(), self.d_dtn.data_ptr())

    def _solvers(self):
        m = self.mesh
        out = []
        for name in ("poisson_vertices", "qg_helmholtz"):
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
                self._select_dt()
                _lib.call("f2d_copy_f64", _lib.ptr(self._glog[i:i + 1]), _lib.ptr(self.d_dt), 1,
                          _lib.stream_ptr())
                self.advance(stats=False)
        if self._graph_key() != key:
            raise RuntimeError("graph capture did not return the state buffers to "
                               "their roles (odd number of buffer swaps)")
        self._graph_solves = [(e, e.calls - c0) for e, c0 in zip(sols, before)]
        for e, k in self._graph_solves:
            e.graph_replayed(0, k)
        self._graphs[key] = g
        return g

def _one(x):
    """The single prognostic array of a one-field prognostic namedtuple."""
    if x is None:
        return None
    return x[0] if len(x) == 1 else x
