"""Restart and checkpoint I/O (restart.py:1-58).

The reference restarts from its netCDF history file: ``read_param``,
``set_state_and_time_from_file`` (the last record of ux, uy, t, ite), then
``integrator.diag``; ``get_model_from_restart`` does both.  The same
functions are here (files through fluids2d_amd.ncfile: netCDF4 when
importable, else NetCDF-3 classic through scipy), and they read whichever
prognostic variables the history stores (``var_to_store``), so eulerpsi
(omega), qg (pv), rsw/qgrsw (u, h) restart too.

``save_checkpoint`` / ``load_checkpoint`` need no NetCDF: an .npz holding
every state array (float64, which holds an FP32 run's values exactly), the
Param (JSON) and the time with its Kahan compensation.  The streamfunction
and velocities are stored with the prognostic fields, so the Poisson solves
resume from the same initial guess: continuing from a checkpoint reproduces
an uninterrupted run bit for bit (FP64; FP32 up to the warm-started direct
solve's row sums, a rounding-level difference).
"""
import json

import numpy as np

from .states import vectors


def get_keys_from_param(param):
    """restart.py:5-9: the names of the parameters a history file stores."""
    return param.parameters()


def read_param(ncfile):
    """restart.py:12-19"""
    from .ncfile import Dataset
    from .param import Param
    param = Param()
    keys = get_keys_from_param(param)
    with Dataset(ncfile) as nc:
        for key in keys:
            if key in nc.ncattrs():
                val = nc.getncattr(key)
                if isinstance(val, str) and val == "None":
                    val = None
                elif isinstance(val, str) and (key == "noslip" or
                                               isinstance(getattr(param, key), (list, tuple))):
                    # a NetCDF-3 file holds a list of names as one string
                    val = [v for v in val.split(",") if v]
                setattr(param, key, val)
    return param


def check_two_params_are_same(param1, param2):
    """restart.py:22-29"""
    keys = get_keys_from_param(param1)
    return {key: np.all(getattr(param1, key) == getattr(param2, key)) for key in keys}


def set_state_and_time_from_file(ncfile, model):
    """restart.py:32-44: the last record of the stored prognostic variables,
    t and ite; then diag."""
    from .ncfile import Dataset
    from .states import get_specs
    s, t = model.state, model.time
    prog = get_specs(model.param).prognostic
    with Dataset(ncfile) as nc:
        for name in prog:
            if name in vectors:
                v = getattr(s, name)
                v.x[:] = nc.variables[name + "x"][-1]
                v.y[:] = nc.variables[name + "y"][-1]
            else:
                getattr(s, name)[:] = nc.variables[name][-1]
        t.t = float(nc.variables["t"][-1])
        t.ite = int(nc.variables["ite"][-1])
    t.ite0 = t.ite
    t.t0 = t.t
    model.integrator.diag(s)


def get_model_from_param(ncfile):
    """restart.py:47-50"""
    from .model import Model
    return Model(read_param(ncfile))


def get_model_from_restart(ncfile):
    """restart.py:53-57"""
    from .model import Model
    model = Model(read_param(ncfile))
    set_state_and_time_from_file(ncfile, model)
    return model


# -- native checkpoints ---------------------------------------------------------------
def _param_json(param):
    atts = {}
    for k in get_keys_from_param(param):
        v = getattr(param, k)
        if isinstance(v, np.ndarray):
            v = v.tolist()
        try:
            json.dumps(v)
        except TypeError:
            v = repr(v)
        atts[k] = v
    return json.dumps(atts)


def save_checkpoint(model, path):
    """Every state array of `model` (copied from the device), its mask, the
    Param and the time, into the .npz file `path`."""
    model.integrator.end(model.state)
    arrays = {}
    for name in model.state._fields:
        v = getattr(model.state, name)
        if hasattr(v, "_fields"):
            for c in v._fields:
                arrays[f"state.{name}.{c}"] = np.asarray(getattr(v, c))
        else:
            arrays[f"state.{name}"] = np.asarray(v)
    arrays["mesh.msk"] = np.asarray(model.mesh.msk)
    t = model.time
    arrays["time"] = np.array([t.t, t.dt, t._c, t.t0], dtype=np.float64)
    arrays["ite"] = np.array([t.ite, t.ite0], dtype=np.int64)
    arrays["param"] = np.array(_param_json(model.param))
    np.savez(path, **arrays)


def restore_checkpoint(model, path):
    """Load a checkpoint into an existing model of the same configuration
    (state, mask, time); the next ``run`` continues from it."""
    with np.load(path, allow_pickle=False) as d:
        msk = d["mesh.msk"]
        if not np.array_equal(msk, model.mesh.msk):
            model.mesh.msk[...] = msk
            model.mesh.finalize()
            model.integrator.model.mesh = model.mesh
        for key in d.files:
            if not key.startswith("state."):
                continue
            parts = key.split(".")
            v = getattr(model.state, parts[1])
            if len(parts) == 3:
                v = getattr(v, parts[2])
            v[...] = d[key]
        tt, ti = d["time"], d["ite"]
        model.time.t, model.time.dt, model.time._c, model.time.t0 = (float(x) for x in tt)
        model.time.ite, model.time.ite0 = int(ti[0]), int(ti[1])
    model.integrator.begin(model.state)


def load_checkpoint(path):
    """A new Model built from the checkpoint's Param, with its state and
    time restored."""
    from .model import Model
    from .param import Param
    import fluids2d_amd.param as fp
    with np.load(path, allow_pickle=False) as d:
        atts = json.loads(str(d["param"]))
    quiet = fp.QUIET
    fp.QUIET = True
    try:
        p = Param()
    finally:
        fp.QUIET = quiet
    known = set(p.parameters())
    for k, v in atts.items():
        if k not in known:
            p.add_parameter(k)
        setattr(p, k, v)
    model = Model(p)
    restore_checkpoint(model, path)
    return model
