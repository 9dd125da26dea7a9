"""Model: driver loop (model.py:13-123) over the GPU integrator.

Same attributes and methods as the reference: ``param, mesh, state,
integrator, time, io, callbacks, diags``, ``run()``, ``step()``,
``set_dt()``, ``add_forcing()``.  ``run()`` keeps the state on the device and
copies it to the host numpy ``state`` before history writes, callbacks and
diagnostics, and at the end of the run.
"""
import signal
from time import time as wallclock

from .integrators import get_integrator
from .io import IO
from .meshes import Mesh
from .states import State
from .timeline import Time


class Model:
    def __init__(self, param):
        param.check()
        self.param = param
        self.mesh = Mesh(param)
        self.state = State(param, self.mesh.shape)
        self.set_integrator()
        self.time = Time(param)
        self.mesh.time = self.time  # forcing functions read mesh.time (forced_convection.py:21)
        self.io = IO(param, self.mesh, self.state, self.time)
        self.callbacks = []
        self.diags = []

    def set_integrator(self):
        self.integrator = get_integrator(self.param, self.mesh, self.state)

    def run(self):
        if self.param.animation:
            raise NotImplementedError("live animation is not part of the GPU path "
                                      "(use nhis/var_to_store history output)")
        self.stop = False

        def signal_handler(sig, frame):
            print('\n hit ctrl-C, stopping', end='')
            self.stop = True

        old = signal.signal(signal.SIGINT, signal_handler)
        try:
            tic = wallclock()
            self.save_to_file()
            integ = self.integrator
            integ.begin(self.state)
            self.nonfinite = []
            while (not self.time.finished) and (not self.stop):
                integ.set_dt(self.time)
                integ.device_step(self.time)
                self.progress()
                self._host_hooks()
            if not self.nonfinite:
                self._check_finite()
            integ.end(self.state)
            self.progress()
            toc = wallclock()
            self.print_perf(toc - tic)
            self.finalize()
        finally:
            signal.signal(signal.SIGINT, old)

    def _host_hooks(self):
        """Points of the loop where host code reads (or edits) the state.
        Device diagnostics (``diag.device``) read the device state: they
        need no host copy."""
        host = [d for d in self.diags if not getattr(d, "device", False)]
        for d in self.diags:
            if getattr(d, "device", False):
                d()
        if not (host or self.time.save_to_file):
            return
        if self._check_finite():
            return
        self.integrator.end(self.state)
        for d in host:
            d()
        self.save_to_file()

    def _check_finite(self):
        """At the loop's synchronisation points (history, host diagnostics,
        the end): stop cleanly, as on ctrl-C, when a prognostic field holds
        a non-finite value on a wet cell (no counterpart in the reference).
        Returns True when it stops the run."""
        bad = self.integrator.nonfinite()
        if not bad:
            return False
        print(f"\n non-finite {', '.join(bad)} at ite={self.time.ite}, stopping", end="")
        self.nonfinite = bad
        self.stop = True
        return True

    def finalize(self):
        for diag in self.diags:
            diag.finalize()

    def step(self, nsteps=1):
        for _ in range(nsteps):
            self.integrator.step(self.state, self.time)

    def set_dt(self):
        """model.py:71-87 (the CFL maximum is taken on the device)."""
        self.integrator.begin(self.state)
        self.integrator.set_dt(self.time)

    def print_perf(self, elapsed):
        print()
        nite = self.time.ite - self.time.ite0
        if nite == 0:
            print(f"Elapsed: {elapsed:.2f} s")
            return
        perf = elapsed / (self.mesh.nx * self.mesh.ny * nite)
        print(f"Elapsed: {elapsed:.2f} s  perf: {perf:.2e} s/dof")

    def progress(self):
        if (self.time.ite % self.param.nprint == 0) or (self.time.finished):
            msg = [f"\rite={self.time.ite}", f"{self.time.tostring()}",
                   f"dt={self.time.dt:.2g}"]
            print(" ".join(msg), end="")

    def save_to_file(self):
        if self.time.save_to_file:
            self.io.write(self.state, self.time)

    def execute_callbacks(self):
        for func in self.callbacks:
            func(self.param, self.mesh, self.state, self.time)

    def compute_diags(self):
        for diag in self.diags:
            diag()

    def add_forcing(self, forcing):
        """model.py:121-123: forcing(param, mesh, s, ds) added after every
        rhs; it receives the device state and tendency (torch tensors, the
        padded layout), see RKIntegrator.add_forcing."""
        self.integrator.add_forcing(forcing)
