"""NetCDF files of the history output, the restart readers and the Bulk
diagnostics.

The reference writes them with netCDF4 (io.py:1, diagnostics.py:1,
restart.py:1).  ``Dataset`` returns a ``netCDF4.Dataset`` when netCDF4 is
importable, so the files are the reference's NETCDF4 files.  Without it the
same calls go to a NetCDF-3 classic file (64-bit offsets) written and read
by ``scipy.io.netcdf_file``, through the part of the netCDF4 interface this
package uses: the context manager, ``setncatts`` / ``ncattrs`` /
``getncattr``, ``createDimension`` (one unlimited), ``createVariable``,
``variables`` (slice assignment that grows the record dimension, reads)
and ``len(dimensions[name])``.

NetCDF-3 attributes are text or numbers.  A list or tuple of strings (such
as ``var_to_store``) is stored as one string, the items joined by commas,
and read back as that string: ``restart.read_param`` splits it again where
the parameter's default is a list.  Floats are stored as doubles, other
values as their ``str``.
"""
import numpy as np


def Dataset(path, mode="r", format="NETCDF4"):
    """netCDF4.Dataset(path, mode, format) when netCDF4 is importable, else a
    NetCDF-3 classic file with the same interface (see the module text)."""
    try:
        import netCDF4
    except ImportError:
        return ClassicDataset(path, mode)
    if mode == "w":
        return netCDF4.Dataset(path, mode, format=format)
    return netCDF4.Dataset(path, mode)


def _to_attribute(v):
    if isinstance(v, (bool, np.bool_)):
        return np.int32(int(v))
    if isinstance(v, (int, np.integer)):
        return np.int32(v) if -2**31 <= int(v) < 2**31 else np.array([float(v)], ">f8")
    if isinstance(v, (float, np.floating)):
        return np.array([float(v)], dtype=">f8")
    if isinstance(v, str):
        return v
    if isinstance(v, (list, tuple, np.ndarray)):
        items = list(v)
        if all(isinstance(x, str) for x in items):
            return ",".join(items)
        try:
            return np.asarray(items, dtype=">f8")
        except (TypeError, ValueError):
            return str(v)
    return str(v)


def _from_attribute(v):
    if isinstance(v, bytes):
        return v.decode("utf-8")
    if isinstance(v, np.ndarray) and v.size == 1:
        return v.reshape(()).item()
    if isinstance(v, np.generic):
        return v.item()
    return v


class _Dimensions:
    def __init__(self, f):
        self._f = f

    def __getitem__(self, name):
        size = self._f.dimensions[name]
        return range(self._f._recs if size is None else size)

    def __contains__(self, name):
        return name in self._f.dimensions

    def keys(self):
        return self._f.dimensions.keys()


class ClassicDataset:
    """A NetCDF-3 classic file through a subset of netCDF4.Dataset."""

    _modes = {"r": "r", "w": "w", "r+": "a", "a": "a"}

    def __init__(self, path, mode="r"):
        from scipy.io import netcdf_file
        if mode not in self._modes:
            raise ValueError(f"unsupported mode {mode!r}")
        self._f = netcdf_file(path, self._modes[mode], mmap=False, version=2)
        self.variables = self._f.variables
        self.dimensions = _Dimensions(self._f)
        self._pending = []  # dimensions not yet written (see createDimension)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        self._flush_dimensions()
        self._f.close()

    def setncatts(self, atts):
        for k, v in atts.items():
            setattr(self._f, k, _to_attribute(v))

    def ncattrs(self):
        return list(self._f._attributes)

    def getncattr(self, name):
        return _from_attribute(self._f._attributes[name])

    def createDimension(self, name, size):
        # NetCDF-3 (scipy) wants the unlimited dimension first; netCDF4
        # callers may create it last (io.py:62-65): dimensions are written
        # when the first variable is created, the unlimited one first
        self._pending.append((name, size))

    def _flush_dimensions(self):
        if not self._pending:
            return
        for name, size in sorted(self._pending, key=lambda d: d[1] is not None):
            self._f.createDimension(name, size)
        self._pending = []

    def createVariable(self, name, dtype, dimensions):
        self._flush_dimensions()
        return self._f.createVariable(name, dtype, dimensions)
