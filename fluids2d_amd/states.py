"""Model states (states.py:1-94): which fields a model carries and how they
are allocated.

Interface kept: ``State(param, shape)`` gives a namedtuple of float64 numpy
arrays of shape (ny+6, nx+6), prognostic fields first, vector fields as
``Vector(x, y)``; ``Prognostic(param, shape)`` the prognostic part only;
``get_specs(param)`` the two name tuples (``.variables``, ``.prognostic``),
with a passive tracer (``param.tracer``) added after the model's prognostic
fields; ``vectors`` the names held as vectors.  The field lists per model are
the reference's (they are the state's interface).

Here each model's fields are stored as two tuples, prognostic and
diagnostic, so a state is prognostic-first by construction.  The GPU
integrator keeps a device twin of the same structure in the compute
precision (``allocate_device``), copied to and from the host at the
synchronisation points the integrator documents.
"""
from collections import namedtuple
from functools import lru_cache

import numpy as np
import torch

#  model:         (prognostic,   diagnostic)
FIELDS = {
    "euler":       (("u",),      ("U", "omega", "ke", "p", "div", "flx")),
    "eulerpsi":    (("omega",),  ("U", "psi", "vomega", "flx")),
    "boussinesq":  (("b", "u"),  ("U", "omega", "ke", "p", "div", "flx")),
    "hydrostatic": (("b", "uh"), ("U", "omega", "ke", "p", "div", "flx")),
    "rsw":         (("u", "h"),  ("U", "omega", "ke", "p", "flx", "pv")),
    "qgrsw":       (("u", "h"),  ("U", "omega", "ke", "p", "flx", "pv", "psi")),
    "qg":          (("pv",),     ("U", "h", "flx", "work", "psi")),
    "advection":   (("q",),      ("U", "flx")),
    "vectoradv":   (("v",),      ("U", "omega", "q")),
}

# fields held as (x, y) pairs on the C-grid faces
vectors = ["u", "U", "flx", "v"]

Specs = namedtuple("specs", ("variables", "prognostic"))
Vector = namedtuple("vector", ("x", "y"))


def get_specs(param):
    """All field names of param.model's state (prognostic first) and the
    prognostic ones; a tracer named by param.tracer joins the prognostic."""
    prog, diag = FIELDS[param.model]
    tracer = getattr(param, "tracer", None)
    if tracer not in (None, "None"):
        prog = prog + (tracer,)
    return Specs(prog + diag, prog)


@lru_cache(maxsize=None)
def _record(typename, names):
    return namedtuple(typename, names)


def _build(typename, names, make):
    return _record(typename, tuple(names))(
        *[Vector(make(), make()) if n in vectors else make() for n in names])


def State(param, shape):
    """Every field of the model, zero-filled float64 host arrays."""
    return _build("state", get_specs(param).variables, lambda: np.zeros(shape))


def Prognostic(param, shape):
    """The prognostic fields only (RK scratch tendencies)."""
    return _build("state", get_specs(param).prognostic, lambda: np.zeros(shape))


# --------------------------------------------------------------------------
# device twins
# --------------------------------------------------------------------------
def allocate_device(variables, shape, dtype, device):
    """Zero-filled device tensors with the structure of a host state."""
    return _build("dstate", variables,
                  lambda: torch.zeros(shape, dtype=dtype, device=device))


def leaves(x):
    """Arrays of a (nested) namedtuple in field order."""
    if hasattr(x, "_fields"):
        out = []
        for v in x:
            out += leaves(v)
        return out
    return [x]


def upload(host, dev, names=None):
    """host numpy state -> device state (in place)."""
    for n in (names or host._fields):
        for h, d in zip(leaves(getattr(host, n)), leaves(getattr(dev, n))):
            d.copy_(torch.from_numpy(np.ascontiguousarray(h)).to(d.device, d.dtype),
                    non_blocking=False)


def download(dev, host, names=None):
    """device state -> host numpy state (in place, float64)."""
    for n in (names or host._fields):
        for h, d in zip(leaves(getattr(host, n)), leaves(getattr(dev, n))):
            h[...] = d.to(torch.float64).cpu().numpy()
