"""tools.set_uv_from_omega (tools.py:6-26): velocity from vertex vorticity
through the vertex Poisson solve and the perpendicular gradient."""
import numpy as np
import torch

from . import operators as op
from .states import Vector


def set_uv_from_omega(model, omega, u, contravariant=False):
    r"""Compute velocity `u` from a distribution of vorticity `omega`
    using u = \nabla^\perp \psi and \nabla^2 \psi = \omega.

    Note: `omega` should be defined on vertices.  Host numpy arrays are
    solved on the device and written back in place.
    """
    mesh = model.mesh
    dt = torch.float64 if mesh.param.precision == "fp64" else torch.float32
    d = mesh.device
    dom = torch.from_numpy(np.ascontiguousarray(omega, dtype=np.float64)).to(d).to(dt)
    psi = torch.zeros_like(dom)
    mesh.poisson_vertices.solve(dom, psi)
    U = Vector(torch.from_numpy(np.ascontiguousarray(u.x)).to(d).to(dt),
               torch.from_numpy(np.ascontiguousarray(u.y)).to(d).to(dt))
    op.perpgrad(mesh, psi, U, contravariant=contravariant, fill=False)
    u.x[...] = U.x.double().cpu().numpy()
    u.y[...] = U.y.double().cpu().numpy()
