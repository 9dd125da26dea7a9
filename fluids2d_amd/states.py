"""State / Prognostic (states.py:1-94).

The user-facing ``State`` is the reference's: a namedtuple of float64 numpy
arrays of shape (ny+6, nx+6), written in place by experiment scripts.  The GPU
integrator keeps a device twin (``device_state``) of the same structure in the
compute precision and copies between them at the synchronisation points
described in the integrator.
"""
from collections import namedtuple

import numpy as np
import torch

Specs = namedtuple("specs", ("variables", "prognostic"))

model_specs = {
    "euler": Specs(("u", "U", "omega", "ke", "p", "div", "flx"), ("u",)),
    "eulerpsi": Specs(("omega", "U", "psi", "vomega", "flx"), ("omega",)),
    "boussinesq": Specs(("b", "u", "U", "omega", "ke", "p", "div", "flx"), ("b", "u")),
    "hydrostatic": Specs(("b", "uh", "U", "omega", "ke", "p", "div", "flx"), ("b", "uh")),
    "rsw": Specs(("u", "h", "U", "omega", "ke", "p", "flx", "pv"), ("u", "h")),
    "qgrsw": Specs(("u", "h", "U", "omega", "ke", "p", "flx", "pv", "psi"), ("u", "h")),
    "qg": Specs(("pv", "U", "h", "flx", "work", "psi"), ("pv",)),
    "advection": Specs(("q", "U", "flx"), ("q",)),
    "vectoradv": Specs(("v", "U", "omega", "q"), ("v",)),
}

vectors = ["u", "U", "flx", "v"]


def get_specs(param):
    specs = model_specs[param.model]
    if (param.tracer is None) | (param.tracer == "None"):
        return specs
    return add_tracer_to_spec(specs, param.tracer)


def add_tracer_to_spec(specs, tracer):
    p = specs.prognostic + (tracer,)
    v = p + specs.variables[len(specs.prognostic):]
    return Specs(v, p)


def State(param, shape):
    specs = get_specs(param)
    check_specs_are_well_ordered(specs)
    return allocate_state(param.model, specs.variables, shape)


def Prognostic(param, shape):
    specs = get_specs(param)
    return allocate_state(param.model, specs.prognostic, shape)


def check_specs_are_well_ordered(specs):
    nprognostic = len(specs.prognostic)
    assert specs.variables[:nprognostic] == specs.prognostic


def define_Namedtuple(name, variables):
    class Namedtuple(namedtuple(name, variables)):
        def __repr__(self):
            return f"{name} with {variables}"
    return Namedtuple


Vector = define_Namedtuple("vector", ("x", "y"))


def allocate_var(name, shape):
    if name in vectors:
        return Vector(x=np.zeros(shape), y=np.zeros(shape))
    return np.zeros(shape)


def allocate_state(name, variables, shape):
    S = define_Namedtuple("state", variables)
    return S(**{var: allocate_var(var, shape) for var in variables})


# --------------------------------------------------------------------------
# device twins
# --------------------------------------------------------------------------
def allocate_device(variables, shape, dtype, device):
    S = define_Namedtuple("dstate", variables)

    def one(name):
        z = lambda: torch.zeros(shape, dtype=dtype, device=device)  # noqa: E731
        return Vector(x=z(), y=z()) if name in vectors else z()

    return S(**{v: one(v) for v in variables})


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
