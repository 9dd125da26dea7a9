"""The NetCDF layer of the history output, restart readers and Bulk
diagnostics (fluids2d_amd.ncfile) without netCDF4: NetCDF-3 classic files
through scipy, used with the calls of io.py:60-92, restart.py:12-44 and
diagnostics.py:79-101.  Parameters round-trip at full precision."""
import numpy as np
import pytest

from fluids2d_amd.ncfile import ClassicDataset, Dataset


def test_history_layout_roundtrip(tmp_path):
    path = str(tmp_path / "history.nc")
    atts = {"model": "eulerpsi", "nx": 48, "cfl": 0.9, "dt": 0.0123456789012345,
            "xperiodic": 1, "noslip": "None", "var_to_store": ["omega", "psi"],
            "Lx": 2.5, "tend": 1e9, "maxite": 123456}
    with ClassicDataset(path, "w") as nc:
        nc.setncatts(atts)
        for dim, size in {"x": 5, "y": 4, "t": None}.items():
            nc.createDimension(dim, size)
        nc.createVariable("xc", "f", ("y", "x"))
        nc.createVariable("t", "f", ("t",))
        nc.createVariable("ite", "i4", ("t",))
        v = nc.createVariable("omega", "f", ("t", "y", "x"))
        v.standard_name = "omega"
        nc.variables["xc"][:, :] = np.arange(20.0).reshape(4, 5)
    rng = np.random.default_rng(0)
    frames = [rng.standard_normal((4, 5)) for _ in range(3)]
    for kt, f in enumerate(frames):
        with ClassicDataset(path, "r+") as nc:
            nc.variables["t"][kt] = 0.5 * kt
            nc.variables["ite"][kt] = 10 * kt
            nc.variables["omega"][kt, :, :] = f
    with ClassicDataset(path, "r") as nc:
        assert len(nc.dimensions["t"]) == 3
        assert len(nc.dimensions["x"]) == 5
        assert set(atts) <= set(nc.ncattrs())
        assert nc.getncattr("model") == "eulerpsi"
        assert nc.getncattr("nx") == 48
        assert nc.getncattr("dt") == 0.0123456789012345  # stored as a double
        assert nc.getncattr("noslip") == "None"
        assert nc.getncattr("var_to_store") == "omega,psi"
        last = np.array(nc.variables["omega"][-1])
        assert np.array_equal(last, frames[-1].astype(np.float32))
        assert int(nc.variables["ite"][-1]) == 20
        assert float(nc.variables["t"][-1]) == 1.0
        assert np.array_equal(np.array(nc.variables["xc"][:]), np.arange(20.0).reshape(4, 5))


def test_dataset_falls_back_without_netcdf4(tmp_path):
    try:
        import netCDF4  # noqa: F401
        pytest.skip("netCDF4 is installed: Dataset returns netCDF4 files")
    except ImportError:
        pass
    with Dataset(str(tmp_path / "d.nc"), "w", format="NETCDF4") as nc:
        assert isinstance(nc, ClassicDataset)
        nc.createDimension("t", None)


def test_read_param_splits_lists(tmp_path):
    """restart.read_param (restart.py:12-19) on a classic file: lists of
    names come back as lists, None as None, numbers at full precision."""
    import fluids2d_amd.param as fp
    from fluids2d_amd.io import param_attributes
    from fluids2d_amd.restart import read_param
    fp.QUIET = True
    p = fp.Param()
    p.model, p.nx, p.ny, p.cfl, p.dtmax = "qgrsw", 96, 40, 0.7, 0.0123456789012345
    p.var_to_store, p.noslip, p.xperiodic = ["u", "h"], ["bottom", "top"], True
    path = str(tmp_path / "h.nc")
    with ClassicDataset(path, "w") as nc:
        nc.setncatts(param_attributes(p))
    q = read_param(path)
    assert (q.model, q.nx, q.ny, q.cfl, q.dtmax) == ("qgrsw", 96, 40, 0.7, 0.0123456789012345)
    assert q.var_to_store == ["u", "h"] and q.noslip == ["bottom", "top"]
    assert q.xperiodic == 1
