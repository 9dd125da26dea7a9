"""The C ABI without a GPU: the header, the Python binding and the built
library agree, every entry point cites the reference interface it replaces,
and the product package refuses to run without the device (no CPU
fallback).  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fluids2d_amd.h")


DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(f2d_\w+)\s*\(", re.M)


def _declared():
    text = open(HEADER).read()
    return text, sorted(set(DECL.findall(text)))


def test_header_matches_binding():
    from fluids2d_amd import _lib
    _, names = _declared()
    assert names, "no declarations found"
    assert sorted(_lib.EXPORTS) == names


def test_every_entry_point_cites_its_reference_or_role():
    text, _ = _declared()
    for m in DECL.finditer(text):
        n, i = m.group(1), m.start(1)
        # the comments between the end of the previous declaration and this one
        ends = [m.end() for m in re.finditer(r"\)\s*;|\}[^;{]*;|#define[^\n]*\n|\{",
                                             text[:i])]
        prev = ends[-1] if ends else 0
        assert re.search(r"\.py:\d+|no counterpart", text[prev:i]), n


def test_library_exports_every_symbol():
    from fluids2d_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIBPATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s+(f2d_\w+)", out))
    _, names = _declared()
    missing = set(names) - exported
    assert not missing, missing
    L = _lib.lib()
    for n in names:
        getattr(L, n)
    assert L.f2d_abi_version() == 8 == _lib.ABI_VERSION


def test_library_is_gfx950_code():
    from fluids2d_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("library not built")
    blob = open(_lib.LIBPATH, "rb").read()
    assert b"gfx950" in blob


def test_product_path_needs_the_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    from fluids2d_amd._lib import F2DError
    fp.QUIET = True
    p = f2d.Param()
    p.model = "eulerpsi"
    with pytest.raises(F2DError, match="needs a GPU"):
        f2d.Model(p)


def test_product_package_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "fluids2d_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                src = open(os.path.join(dirpath, f), errors="replace").read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", src, re.M), f
                assert "liboracle" not in src, f


def test_package_api_surface():
    """The drop-in classes carry the reference's methods (checked without a
    GPU: import and attribute lookup only)."""
    import importlib
    for mod in ("fluids2d_amd", "fluids2d_amd.model", "fluids2d_amd.integrators",
                "fluids2d_amd.elliptic", "fluids2d_amd.meshes", "fluids2d_amd.equations",
                "fluids2d_amd.rsw", "fluids2d_amd.adjoint", "fluids2d_amd.ensemble",
                "fluids2d_amd.autotune", "fluids2d_amd.diagnostics", "fluids2d_amd.operators",
                "fluids2d_amd.roofline"):
        importlib.import_module(mod)
    from fluids2d_amd.elliptic import Poisson2D
    from fluids2d_amd.model import Model
    from fluids2d_amd.integrators import RKIntegrator
    from fluids2d_amd.diagnostics import Bulk
    from fluids2d_amd.ensemble import Ensemble
    expect = {
        Poisson2D: ("solve", "solve_c2v", "handle", "read_stats", "reset_stats"),  # elliptic.py:71-87
        Model: ("run", "step", "set_dt", "add_forcing", "compute_diags", "finalize"),  # model.py
        RKIntegrator: ("step", "advance", "device_steps", "begin", "end"),  # integrators.py:63-79
        Bulk: ("__call__", "finalize", "write", "read"),  # diagnostics.py:8-76
        Ensemble: ("steps", "run", "end", "tune_streams"),
    }
    for cls, names in expect.items():
        for n in names:
            assert callable(getattr(cls, n, None)), (cls.__name__, n)
