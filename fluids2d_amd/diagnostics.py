"""Bulk diagnostics (diagnostics.py:8-104) on the device.

Same interface as the reference: ``Bulk(model)``, appended to
``model.diags``; called every step, it records every third iteration the
domain averages ke, enstrophy, vorticity and angular momentum
(diagnostics.py:45-66), buffers ``ndiags`` records and writes them to
``bulk.nc`` (``finalize``/``write``, ``read``).

The averages are FP64 sums over the padded arrays, formed on the device from
the model's device state (`f2d_bulk_diagnostics`, also in FP32 runs), so a
run with diagnostics never copies the state to the host: records stay in
device memory until ``write``.  eulerpsi, whose state has no ke, takes the
cell-centred 1/2 u.U of its velocity.  The netCDF file is written with
netCDF4 when importable, else as NetCDF-3 classic through scipy
(fluids2d_amd.ncfile); ``ncfile=None`` keeps the records in memory only.
"""
from collections import namedtuple

import numpy as np
import torch

from . import _lib

VARIABLES = ("time", "ke", "ens", "vort", "angular")
BulkData = namedtuple("bulk", VARIABLES)


class Bulk:
    device = True  # reads the device state: no host copy per record

    def __init__(self, model, ncfile="bulk.nc", ndiags=1000):
        self.model = model
        s = model.integrator.dstate
        if not hasattr(s, "omega"):
            raise NotImplementedError(f"Bulk needs omega: not in the {model.param.model} state")
        self.ndiags = ndiags
        dev = model.mesh.device
        self.rec = torch.zeros((ndiags, 4), dtype=torch.float64, device=dev)
        self.t = np.zeros(ndiags)
        n = (model.mesh.ny + 6) * (model.mesh.nx + 6)
        self.nparts = min(1024, (n + 255) // 256)
        self.part = torch.zeros(6 * self.nparts, dtype=torch.float64, device=dev)
        self.kt = 0
        self.history = [np.zeros(0) for _ in VARIABLES]
        self.ncfile = ncfile
        self.k0 = 0
        if ncfile is not None:
            self.k0 = get_number_of_records(ncfile)
            if self.k0 == 0:
                create_file(ncfile, model.param, VARIABLES)

    def measure(self, out):
        """The four averages of the current device state into out[0:4]."""
        m, s = self.model.mesh, self.model.integrator.dstate
        # the fields a fused stage does not store between steps
        self.model.integrator.model.materialize(s, names=("U", "ke", "omega"))
        ke = getattr(s, "ke", None)
        _lib.call("f2d_bulk_diagnostics", _lib.dtype_code(s.omega.dtype), m.grid_ref(),
                  _lib.ptr(ke), _lib.ptr(s.omega), _lib.ptr(s.U.x), _lib.ptr(s.U.y),
                  _lib.ptr(m.d_msk), _lib.ptr(self.part), self.nparts, _lib.ptr(out),
                  _lib.stream_ptr())

    def __call__(self):
        if self.model.time.ite % 3 > 0:
            return
        self.measure(self.rec[self.kt])
        self.t[self.kt] = self.model.time.t
        self.kt += 1
        if self.kt == self.ndiags:
            self.write()

    def finalize(self):
        self.write()

    def write(self):
        n = self.kt
        if n == 0:
            return
        rec = self.rec[:n].cpu().numpy()
        cols = [self.t[:n].copy()] + [rec[:, q] for q in range(4)]
        self.history = [np.concatenate([h, c]) for h, c in zip(self.history, cols)]
        if self.ncfile is not None:
            from .ncfile import Dataset
            idx = slice(self.k0, self.k0 + n)
            with Dataset(self.ncfile, "r+") as nc:
                for name, x in zip(VARIABLES, cols):
                    nc.variables[name][idx] = x
        self.k0 += n
        self.kt = 0

    def read(self):
        if self.ncfile is not None:
            from .ncfile import Dataset
            with Dataset(self.ncfile, "r") as nc:
                return BulkData(*[nc.variables[name][:] for name in VARIABLES])
        pending = self.rec[:self.kt].cpu().numpy()
        cols = [self.t[:self.kt]] + [pending[:, q] for q in range(4)]
        return BulkData(*[np.concatenate([h, c]) for h, c in zip(self.history, cols)])


def create_file(ncfile, param, variables):
    """diagnostics.py:79-92"""
    from .ncfile import Dataset
    from .io import param_attributes
    with Dataset(ncfile, "w", format="NETCDF4") as nc:
        nc.setncatts(param_attributes(param))
        nc.createDimension("t", None)
        for name in variables:
            v = nc.createVariable(name, "f", ("t",))
            v.standard_name = name


def get_number_of_records(ncfile):
    """diagnostics.py:95-101"""
    import os
    if not os.path.isfile(ncfile):
        return 0
    from .ncfile import Dataset
    with Dataset(ncfile) as nc:
        return len(nc.dimensions["t"])
