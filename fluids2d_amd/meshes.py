"""Mesh: Cartesian C-grid with a 3-cell halo, masks and stencil orders
(meshes.py:7-186, noslip.py:4-36).

The integer mask logic runs in the HIP library (f2d_set_masks, f2d_set_order,
f2d_slipcoef); the results are kept both on the device (used by the kernels)
and as host numpy arrays with the reference's names, so user scripts that read
``mesh.mskv`` or edit ``mesh.msk`` and call ``finalize()`` work unchanged.
"""
from collections import namedtuple

import numpy as np
import torch

from . import _lib
from .elliptic import Poisson2D

Stencil = namedtuple("stencil", ("x", "y"))


def device_of(param):
    if not torch.cuda.is_available():
        raise _lib.F2DError("fluids2d_amd needs a GPU (torch.cuda is unavailable)")
    return torch.device("cuda", int(getattr(param, "device", 0) or 0))


class Mesh:
    """Mesh: class to handle Cartesian meshes (meshes.py:7-52).

    If the default `msk` is changed then use `mesh.finalize()` to update all
    the derived information: mskx, msky, mskv, the stencil orders and the
    Poisson operators.
    """

    def __init__(self, param):
        self.param = param
        self.device = device_of(param)
        self.shape = get_shape(param)
        self.nx = param.nx
        self.ny = param.ny
        self.dx = param.Lx / self.nx
        self.dy = param.Ly / self.ny
        self.area = self.dx * self.dy
        self.xshift = 1
        self.yshift = self.shape[-1]
        self.grid = _lib.make_grid(param)
        self.set_default_mask()
        self.finalize()

    # -- reference API ----------------------------------------------------
    def finalize(self):
        # bumped on every finalize: captured step graphs hold the solvers
        # and mask arrays of one version
        self.version = getattr(self, "version", 0) + 1
        self._check_mask()
        self.set_masks()
        self.slipcoef = self._slipcoef()
        self.set_stencils(self.param.maxorder)
        # the Neumann "c" operator of the euler pressure projection
        self.poisson_centers = Poisson2D(self, "c") if self.param.model == "euler" else None
        # the vertex solvers are set up on first use (rsw uses neither):
        # attribute access builds them from this finalize's masks
        for name in self._LAZY:
            self.__dict__.pop(name, None)
        self._factories = {"poisson_vertices": lambda: Poisson2D(self, "v")}
        if self.param.model in ["qg", "qgrsw", "rsw"]:
            f0, H, g = self.param.f0, self.param.H, self.param.g
            maindiag = self.area * f0 ** 2 / (g * H)
            self._factories["qg_helmholtz"] = lambda: Poisson2D(self, "v", maindiag=maindiag)
            self.hb = 0
            self.qgcoef = f0 / +H

    _LAZY = ("poisson_vertices", "qg_helmholtz")

    def __getattr__(self, name):
        make = self.__dict__.get("_factories", {}).get(name)
        if make is None:
            raise AttributeError(f"{type(self).__name__!r} object has no attribute {name!r}")
        solver = make()
        setattr(self, name, solver)
        return solver

    def _allocate(self):
        return np.zeros(self.shape, dtype="i1")

    def x(self, which):
        idx = np.arange(self.nx + 2 * self.param.halowidth) - self.param.halowidth
        shift = 0.5 if which in ["c", "y"] else 0.0
        return (idx + shift) * self.dx

    def y(self, which):
        idx = np.arange(self.ny + 2 * self.param.halowidth) - self.param.halowidth
        shift = 0.5 if which in ["c", "x"] else 0.0
        return (idx + shift) * self.dy

    def xy(self, which="c"):
        return np.meshgrid(self.x(which), self.y(which))

    def set_default_mask(self):
        nh = self.param.halowidth
        self.msk = self._allocate()
        xdx = slice(None) if self.param.xperiodic else slice(nh, -nh)
        ydx = slice(None) if self.param.yperiodic else slice(nh, -nh)
        self.msk[ydx, xdx] = 1

    def set_masks(self):
        """meshes.py:77-86, computed on the device."""
        d = self.device
        self.d_msk = torch.from_numpy(np.ascontiguousarray(self.msk, dtype=np.int8)).to(d)
        self.d_mskx = torch.zeros_like(self.d_msk)
        self.d_msky = torch.zeros_like(self.d_msk)
        self.d_mskv = torch.zeros_like(self.d_msk)
        ny, nx = self.shape
        _lib.call("f2d_set_masks", _lib.ptr(self.d_msk), ny, nx, _lib.ptr(self.d_mskx),
                  _lib.ptr(self.d_msky), _lib.ptr(self.d_mskv), _lib.stream_ptr())
        self.mskx = self.d_mskx.cpu().numpy()
        self.msky = self.d_msky.cpu().numpy()
        self.mskv = self.d_mskv.cpu().numpy()

    def _slipcoef(self):
        """get_slipcoef (noslip.py:4-36), computed on the device."""
        ns = self.param.noslip
        if (ns is None) or (ns is False):
            mode, sides = 0, 0
        elif ns is True:
            mode, sides = 1, 0
        else:
            mode = 2
            sides = sum(b for s, b in (("left", 1), ("right", 2), ("bottom", 4),
                                       ("top", 8)) if s in ns)
        ny, nx = self.shape
        self.d_slipcoef = torch.zeros(self.shape, dtype=torch.float64, device=self.device)
        _lib.call("f2d_slipcoef", _lib.ptr(self.d_msk), ny, nx, mode, sides,
                  self.param.halowidth, _lib.ptr(self.d_slipcoef), _lib.stream_ptr())
        return self.d_slipcoef.cpu().numpy()

    def _order(self, m, shift, maxorder):
        o = torch.zeros(self.shape, dtype=torch.int8, device=self.device)
        _lib.call("f2d_set_order", _lib.ptr(m), m.numel(), shift, maxorder, _lib.ptr(o),
                  _lib.stream_ptr())
        return o

    def set_stencils(self, maxorder):
        """meshes.py:88-104, computed on the device."""
        xs, ys = self.xshift, self.yshift
        self.d_oc = Stencil(self._order(self.d_msk, xs, maxorder),
                            self._order(self.d_msk, ys, maxorder))
        mv = (self.d_slipcoef > 0).to(torch.int8)
        ovx = self._order(mv, -xs, maxorder) * self.d_msky
        ovy = self._order(mv, -ys, maxorder) * self.d_mskx
        self.d_ov = Stencil(ovx.to(torch.int8), ovy.to(torch.int8))
        self.d_ok = Stencil(self._order(self.d_mskx, -xs, maxorder),
                            self._order(self.d_msky, -ys, maxorder))
        self.oc = Stencil(*(o.cpu().numpy() for o in self.d_oc))
        self.ov = Stencil(*(o.cpu().numpy() for o in self.d_ov))
        self.ok = Stencil(*(o.cpu().numpy() for o in self.d_ok))

    def fill(self, variable):
        """meshes.py:106-111 on host numpy arrays or device tensors."""
        if hasattr(variable, "_fields"):
            for v in variable:
                self.fill(v)
        elif isinstance(variable, torch.Tensor):
            _lib.call("f2d_fill_halo", _lib.dtype_code(variable.dtype),
                      self.grid_ref(), _lib.ptr(variable), _lib.stream_ptr())
        else:
            fill_halo_array(self.param, variable)

    # -- helpers ---------------------------------------------------------------
    def grid_ref(self):
        import ctypes
        return ctypes.byref(self.grid)

    def _check_mask(self):
        m = self.msk
        assert m.shape == self.shape, "msk has the wrong shape"
        assert np.all((m == 0) | (m == 1)), "msk must hold 0/1"
        nh = self.param.halowidth
        if not self.param.xperiodic and (m[:, :nh].any() or m[:, -nh:].any()):
            raise ValueError("wet halo columns in a closed x direction are not supported")
        if not self.param.yperiodic and (m[:nh, :].any() or m[-nh:, :].any()):
            raise ValueError("wet halo rows in a closed y direction are not supported")


def get_shape(param):
    """meshes.py:122-132"""
    n1 = param.nx + 2 * param.halowidth
    n2 = param.ny + 2 * param.halowidth
    return (n2, n1)


def fill_halo_array(param, array):
    """meshes.py:135-143, plus the y-periodic extension."""
    n = param.halowidth
    if param.xperiodic:
        if array.ndim == 2:
            array[:, :n] = array[:, -2 * n:-n]
            array[:, -n:] = array[:, n:2 * n]
        elif array.ndim == 1:
            array[:n] = array[-2 * n:-n]
            array[-n:] = array[n:2 * n]
    if getattr(param, "yperiodic", False) and array.ndim == 2:
        array[:n, :] = array[-2 * n:-n, :]
        array[-n:, :] = array[n:2 * n, :]
