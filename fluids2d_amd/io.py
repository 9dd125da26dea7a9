"""NetCDF history output (the reference's IO, io.py:1-92).

The file layout is the reference's, since post-processing and the restart
read it:

* dimensions ``x`` (nx+6), ``y`` (ny+6) and the unlimited ``t``;
* ``xc``, ``yc``: cell-centre coordinates, float32 on (y, x);
* ``t`` and ``dt`` float32, ``ite`` int32, on (t,);
* one float32 variable on (t, y, x) per stored field of ``var_to_store``,
  a vector field as two (``ux``, ``uy``), each with a ``standard_name``
  attribute equal to its variable name;
* every parameter as a global attribute, booleans as integers and None as
  the string "None".

``IO(param, mesh, state, time)`` creates the file when ``nhis > 0``; each
``write(state, time)`` appends one record.  Here the writer works out once,
at creation, which arrays each record copies (``HistoryWriter.columns``).
Files go through fluids2d_amd.ncfile: NETCDF4 with netCDF4 importable,
NetCDF-3 classic through scipy otherwise.
"""
from .states import vectors

FIELD_TYPE = "f"  # float32, as the reference stores every field


def param_attributes(param):
    """The global attributes of a history file: every declared parameter,
    booleans as 0/1 and None as "None" (NetCDF has neither)."""
    names = param.parameters() if hasattr(param, "parameters") else \
        [n for n in vars(param) if not n.startswith("_")]
    out = {}
    for name in names:
        v = getattr(param, name)
        if isinstance(v, bool):
            v = int(v)
        out[name] = "None" if v is None else v
    return out


def field_columns(names):
    """(variable name, field name, component) for each stored field: a
    vector field gives its x and y components."""
    cols = []
    for name in names:
        if name in vectors:
            cols += [(name + "x", name, "x"), (name + "y", name, "y")]
        else:
            cols.append((name, name, None))
    return cols


class HistoryWriter:
    """Creates the history file of `param` and appends records to it."""

    def __init__(self, param, mesh):
        self.path = param.outputfile
        self.columns = field_columns(param.var_to_store)
        self.records = 0
        self._create(param, mesh)

    def _create(self, param, mesh):
        from .ncfile import Dataset
        ny, nx = mesh.shape
        with Dataset(self.path, "w", format="NETCDF4") as nc:
            nc.setncatts(param_attributes(param))
            nc.createDimension("x", nx)
            nc.createDimension("y", ny)
            nc.createDimension("t", None)
            xc, yc = mesh.xy()
            for name, coord in (("xc", xc), ("yc", yc)):
                nc.createVariable(name, FIELD_TYPE, ("y", "x"))[:, :] = coord
            for name, typ in (("t", FIELD_TYPE), ("ite", "i4"), ("dt", FIELD_TYPE)):
                nc.createVariable(name, typ, ("t",))
            for var, _, _ in self.columns:
                nc.createVariable(var, FIELD_TYPE, ("t", "y", "x")).standard_name = var

    def append(self, state, time):
        from .ncfile import Dataset
        k = self.records
        with Dataset(self.path, "r+") as nc:
            v = nc.variables
            v["t"][k] = time.t
            v["ite"][k] = time.ite
            v["dt"][k] = time.dt
            for var, field, comp in self.columns:
                a = getattr(state, field)
                v[var][k, :, :] = a if comp is None else getattr(a, comp)
        self.records += 1


class IO:
    """The reference's IO object (io.py:5-23): a history file when nhis > 0."""

    def __init__(self, param, mesh, state, time):
        self.param = param
        self.writer = HistoryWriter(param, mesh) if param.nhis > 0 else None

    @property
    def kt(self):
        """records written so far"""
        return self.writer.records if self.writer is not None else 0

    def write(self, state, time):
        if self.writer is None:
            raise RuntimeError("no history file: param.nhis is 0")
        self.writer.append(state, time)
