"""ctypes binding of libfluids2d_amd.so (the C ABI of include/fluids2d_amd.h).

The library is the only compute path of the package: if it is missing or the
device is unavailable every entry point raises, there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libfluids2d_amd.so")

F64, F32 = 0, 1
METHODS = {"weno": 0, "upwind": 1, "centered": 2, "cweno": 3}
INTEGRATORS = {"rk3": 0, "ef": 1, "enrk3": 2}
NSTAGES = {"rk3": 3, "ef": 1, "enrk3": 3}
POISSON_F64, POISSON_F32_REFINE, POISSON_F32 = 0, 1, 2

EXPORTS = [
    "f2d_abi_version", "f2d_last_error", "f2d_device_synchronize",
    "f2d_set_masks", "f2d_set_order", "f2d_slipcoef", "f2d_gindex",
    "f2d_compflux", "f2d_vortexforce", "f2d_innerproduct",
    "f2d_transport_stage", "f2d_transport_stage_log", "f2d_transport_stage_psi", "f2d_rk_addto", "f2d_addto", "f2d_c2v", "f2d_perpgrad",
    "f2d_fill_halo", "f2d_cfl_dt", "f2d_set_scalar", "f2d_copy_f64",
    "f2d_reserve_stream_scratch",
    "f2d_poisson_create", "f2d_poisson_destroy", "f2d_poisson_solve",
    "f2d_poisson_solve_c2v", "f2d_poisson_stats", "f2d_poisson_reset_stats",
    "f2d_poisson_bench_kernel", "f2d_perpgrad_ex", "f2d_poisson_read_stats",
    "f2d_poisson_method", "f2d_coriolis", "f2d_addgrad", "f2d_sharp", "f2d_vorticity",
    "f2d_ke_finish", "f2d_pressure", "f2d_div", "f2d_c2v_scaled", "f2d_v2c_scaled",
    "f2d_adj_stage", "f2d_adj_combine", "f2d_bulk_diagnostics",
    "f2d_perpgrad_cfl_noalloc", "f2d_sw_momentum", "f2d_sw_momentum_u", "f2d_sw_diag",
    "f2d_qg_pv", "f2d_qg_project", "f2d_transport_stage_u",
    "f2d_poisson_batch_create", "f2d_poisson_batch_solve", "f2d_poisson_batch_destroy",
    "f2d_poisson_set_refine", "f2d_nonfinite",
    "f2d_record_begin", "f2d_record_end", "f2d_recording_size", "f2d_recording_destroy",
    "f2d_launch_batch_create", "f2d_launch_batch_run", "f2d_launch_batch_destroy",
    "f2d_transport_stage_cfl", "f2d_cfl_partials", "f2d_perpgrad_part", "f2d_cfl_finish",
    "f2d_cfl_dt_part",
]


class F2DError(RuntimeError):
    pass


class Grid(ctypes.Structure):
    _fields_ = [("ny", ctypes.c_int64), ("nx", ctypes.c_int64),
                ("halowidth", ctypes.c_int32), ("xperiodic", ctypes.c_int32),
                ("yperiodic", ctypes.c_int32), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("tile", ctypes.c_int32)]


TILES = {0: (16, 64), 1: (8, 64), 2: (16, 32), 3: (32, 32), 4: (8, 128)}


class PoissonOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("nu", ctypes.c_int32),
                ("coarse_sweeps", ctypes.c_int32),
                ("refine_inner", ctypes.c_int32),
                ("tail_points", ctypes.c_int32),
                ("tail_threads", ctypes.c_int32),
                ("nu_coarse", ctypes.c_int32),
                ("nu_tail", ctypes.c_int32),
                ("solver", ctypes.c_int32),
                ("fft_refine", ctypes.c_int32),
                ("fft_warm", ctypes.c_int32),
                ("neumann", ctypes.c_int32)]


SOLVERS = {"auto": 0, "mg": 1, "fft": 2, "cap": 4}


class SwStage(ctypes.Structure):
    """f2d_sw_stage: the RK stage terms of the fused shallow-water kernels,
    fields 0 u.x, 1 u.y, 2 h."""
    _fields_ = [("integrator", ctypes.c_int32), ("stage", ctypes.c_int32),
                ("dt", ctypes.c_void_p), ("prev", (ctypes.c_void_p * 3) * 2),
                ("out", ctypes.c_void_p * 3), ("y", ctypes.c_void_p * 3),
                ("y_in", ctypes.c_void_p * 3), ("dt_log", ctypes.c_void_p)]


def sw_stage(integrator, stage, dt, prev=(), out=(None, None, None), y=(None, None, None),
             dt_log=None, y_in=(None, None, None)):
    """An SwStage from tensors (None for absent arrays); integrator None:
    tendency only.  prev: the earlier stages' (u.x, u.y, h) tendencies;
    y_in: the arrays y is computed from (ping-pong; None: y in place);
    dt_log: a one-element float64 tensor the momentum kernel copies dt to."""
    def a(t):
        return None if t is None else t.data_ptr()
    st = SwStage()
    st.integrator = -1 if integrator is None else INTEGRATORS[integrator]
    st.stage = stage
    st.dt = a(dt)
    for k, fields in enumerate(prev):
        for f, t in enumerate(fields):
            st.prev[k][f] = a(t)
    for f in range(3):
        st.out[f] = a(out[f])
        st.y[f] = a(y[f])
        st.y_in[f] = a(y_in[f])
    st.dt_log = a(dt_log)
    return st


ABI_VERSION = 8  # f2d_abi_version() of the library these bindings describe


_lib = None


def lib():
    """Load the library (raises F2DError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise F2DError(f"{LIBPATH} is not built: run `python "
                       f"fluids2d_amd/csrc/build.py` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIBPATH)
    vp, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    gp = ctypes.POINTER(Grid)
    sigs = {
        "f2d_abi_version": (i32, []),
        "f2d_last_error": (ctypes.c_char_p, []),
        "f2d_device_synchronize": (i32, []),
        "f2d_set_masks": (i32, [vp, i64, i64, vp, vp, vp, vp]),
        "f2d_set_order": (i32, [vp, i64, i64, i32, vp, vp]),
        "f2d_slipcoef": (i32, [vp, i64, i64, i32, i32, i32, vp, vp]),
        "f2d_gindex": (i32, [vp, i64, vp, ctypes.POINTER(i64), vp]),
        "f2d_compflux": (i32, [i32, vp, vp, vp, vp, i64, i64, i32, vp]),
        "f2d_vortexforce": (i32, [i32, vp, vp, vp, vp, i64, i64, i64, i32, i32, vp]),
        "f2d_innerproduct": (i32, [i32, vp, vp, vp, vp, i64, i64, i32, vp]),
        "f2d_transport_stage": (i32, [i32, gp, i32, vp, vp, vp, vp, vp, vp, vp,
                                      vp, vp, vp, i32, i32, vp, vp]),
        "f2d_transport_stage_log": (i32, [i32, gp, i32, vp, vp, vp, vp, vp, vp, vp,
                                          vp, vp, vp, i32, i32, vp, vp, vp]),
        "f2d_transport_stage_cfl": (i32, [i32, gp, i32, vp, vp, vp, vp, vp, vp, vp,
                                          vp, vp, vp, i32, i32, vp, vp, i64, d, d, vp, vp]),
        "f2d_cfl_partials": (i32, [gp, ctypes.POINTER(i64)]),
        "f2d_perpgrad_part": (i32, [i32, gp, vp, vp, vp, vp, vp, d, d, i32, vp, i64, vp]),
        "f2d_cfl_finish": (i32, [vp, i64, vp, vp, d, d, vp]),
        "f2d_transport_stage_psi": (i32, [i32, gp, i32, vp, vp, vp, vp, vp, d, d, i32, vp,
                                          vp, vp, vp, vp, vp, i32, i32, vp, vp]),
        "f2d_rk_addto": (i32, [i32, vp, vp, vp, vp, i64, i32, i32, vp, vp]),
        "f2d_addto": (i32, [i32, vp, vp, vp, vp, i64, i32, d, d, d, vp, i32, vp]),
        "f2d_c2v": (i32, [i32, gp, vp, vp, vp, i32, vp]),
        "f2d_perpgrad": (i32, [i32, gp, vp, vp, vp, vp, vp, i32, vp, vp]),
        "f2d_perpgrad_ex": (i32, [i32, gp, vp, vp, vp, vp, vp, d, d, i32, vp, vp,
                                  d, d, vp]),
        "f2d_perpgrad_cfl_noalloc": (i32, [i32, gp, vp, vp, vp, vp, vp, d, d, i32, vp, vp,
                                           d, d, vp]),
        "f2d_fill_halo": (i32, [i32, gp, vp, vp]),
        "f2d_cfl_dt": (i32, [i32, vp, vp, i64, d, d, vp, vp]),
        "f2d_cfl_dt_part": (i32, [i32, vp, vp, i64, d, d, vp, i64, vp, vp]),
        "f2d_set_scalar": (i32, [vp, d, vp]),
        "f2d_copy_f64": (i32, [vp, vp, i32, vp]),
        "f2d_reserve_stream_scratch": (i32, [vp]),
        "f2d_poisson_create": (i32, [vp, gp, d, ctypes.POINTER(PoissonOpts),
                                     ctypes.POINTER(vp)]),
        "f2d_poisson_destroy": (i32, [vp]),
        "f2d_poisson_method": (i32, [vp]),
        "f2d_coriolis": (i32, [i32, gp, vp, vp, vp, vp, d, vp]),
        "f2d_addgrad": (i32, [i32, gp, vp, vp, vp, vp, vp, vp]),
        "f2d_sharp": (i32, [i32, gp, vp, vp, vp, vp, vp]),
        "f2d_vorticity": (i32, [i32, gp, vp, vp, vp, vp, vp]),
        "f2d_ke_finish": (i32, [i32, gp, vp, vp, vp]),
        "f2d_pressure": (i32, [i32, gp, vp, vp, d, d, vp]),
        "f2d_div": (i32, [i32, gp, vp, vp, vp, vp, vp]),
        "f2d_c2v_scaled": (i32, [i32, gp, vp, d, vp, vp, i32, vp]),
        "f2d_v2c_scaled": (i32, [i32, gp, vp, d, vp, vp, vp, vp]),
        "f2d_adj_stage": (i32, [i32, gp, vp, vp, vp, vp, vp, vp, d, d, d, vp, vp, vp, vp,
                                vp, vp]),
        "f2d_adj_combine": (i32, [i32, gp, vp, vp, vp, vp, vp]),
        "f2d_bulk_diagnostics": (i32, [i32, gp, vp, vp, vp, vp, vp, vp, i32, vp, vp]),
        "f2d_sw_momentum": (i32, [i32, gp, i32, vp, vp, vp, vp, vp, d, vp, vp, vp, vp, i32,
                                  ctypes.POINTER(SwStage), vp]),
        "f2d_sw_momentum_u": (i32, [i32, gp, i32, vp, vp, vp, vp, vp, d, vp, vp, d, d, vp, vp,
                                    i32, ctypes.POINTER(SwStage), vp]),
        "f2d_transport_stage_u": (i32, [i32, gp, i32, vp, vp, vp, vp, vp, vp, vp,
                                        vp, vp, vp, i32, i32, vp, vp]),
        "f2d_sw_diag": (i32, [i32, gp, i32, vp, vp, vp, vp, vp, vp, vp, d, d, vp, vp, vp, vp,
                              vp, vp]),
        "f2d_qg_pv": (i32, [i32, gp, vp, vp, vp, vp, vp, d, vp, vp]),
        "f2d_qg_project": (i32, [i32, gp, vp, vp, vp, vp, vp, d, vp, vp, vp,
                                 ctypes.POINTER(SwStage), vp]),
        "f2d_poisson_solve": (i32, [vp, vp, vp, d, i32, ctypes.POINTER(i32),
                                    ctypes.POINTER(d), vp]),
        "f2d_poisson_solve_c2v": (i32, [vp, vp, vp, d, i32, ctypes.POINTER(i32),
                                        ctypes.POINTER(d), vp]),
        "f2d_poisson_stats": (vp, [vp]),
        "f2d_poisson_reset_stats": (i32, [vp, vp]),
        "f2d_poisson_read_stats": (i32, [vp, ctypes.POINTER(d), vp]),
        "f2d_poisson_bench_kernel": (i32, [vp, i32, i32, ctypes.POINTER(d),
                                           ctypes.POINTER(d), vp]),
        "f2d_poisson_batch_create": (i32, [ctypes.POINTER(vp), i32, ctypes.POINTER(vp),
                                           ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "f2d_poisson_batch_solve": (i32, [vp, i32, vp]),
        "f2d_poisson_batch_destroy": (i32, [vp]),
        "f2d_poisson_set_refine": (i32, [vp, i32]),
        "f2d_nonfinite": (i32, [i32, gp, vp, vp, vp, vp]),
        "f2d_record_begin": (i32, []),
        "f2d_record_end": (i32, [ctypes.POINTER(vp)]),
        "f2d_recording_size": (i32, [vp]),
        "f2d_recording_destroy": (i32, [vp]),
        "f2d_launch_batch_create": (i32, [ctypes.POINTER(vp), i32, ctypes.POINTER(vp)]),
        "f2d_launch_batch_run": (i32, [vp, vp]),
        "f2d_launch_batch_destroy": (i32, [vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.f2d_abi_version() != ABI_VERSION:
        raise F2DError(f"{LIBPATH} has ABI version {L.f2d_abi_version()}, this package "
                       f"binds version {ABI_VERSION}: rebuild it")
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().f2d_last_error().decode(errors="replace")
        raise F2DError(f"fluids2d_amd: {msg} (code {rc})")


def call(name, *args):
    check(getattr(lib(), name)(*args))


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class capture:
    """``with capture(graph):`` -- torch.cuda.graph with Python's cyclic
    garbage collector held off: a collection during the capture could run a
    finaliser that frees device memory (a solver handle), which a capture in
    global mode does not allow."""

    def __init__(self, graph, **kw):
        import torch
        self._ctx = torch.cuda.graph(graph, **kw)

    def __enter__(self):
        import gc
        self._gc = gc.isenabled()
        gc.disable()
        return self._ctx.__enter__()

    def __exit__(self, *exc):
        import gc
        try:
            return self._ctx.__exit__(*exc)
        finally:
            if self._gc:
                gc.enable()


def capturing():
    """True while the current stream is being captured into a graph."""
    try:
        import torch
        return torch.cuda.is_current_stream_capturing()
    except Exception:  # also at interpreter shutdown
        return False


def dtype_code(dt):
    import torch
    if dt == torch.float64:
        return F64
    if dt == torch.float32:
        return F32
    raise F2DError(f"unsupported dtype {dt}")


def make_grid(param):
    g = Grid()
    g.ny, g.nx = param.ny, param.nx
    g.halowidth = param.halowidth
    g.xperiodic = int(bool(param.xperiodic))
    g.yperiodic = int(bool(getattr(param, "yperiodic", False)))
    g.dx = param.Lx / param.nx
    g.dy = param.Ly / param.ny
    g.tile = int(getattr(param, "tile", 0) or 0)
    return g
