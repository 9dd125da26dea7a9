"""Param: the flat configuration object of the reference (param.py:10-109).

Interface kept: the field names and defaults, the accepted values of the
string fields, ``check()`` failing with AssertionError on a value it does not
accept or on an attribute that is no parameter, ``add_parameter(name)`` to
declare one more, and the help text printed by the constructor.

Here the fields are one table, ``SCHEMA``: each row gives a name, its
default, the values ``check()`` accepts (None: any) and a line of help.  The
constructor copies the defaults; ``parameters()`` lists the declared names
(the table's plus those added), which is what the history file stores and
the restart reads back.

Fields of the MI355X path (after the reference's in the table):

* ``precision``      "fp64" (reproduces the reference) or "fp32" (FP32
                     stencils and solves, FP64 iterative refinement of the
                     Poisson residual)
* ``poisson_rtol``   relative residual of the Poisson solves (None: 1e-13 in
                     FP64, 1e-9 for the refined FP32 mode)
* ``poisson_maxit``  iteration cap of the Poisson solves (None: 100 in FP64,
                     8 refinements in FP32)
* ``poisson_nu``     red-black Gauss-Seidel sweeps before/after each level
* ``poisson_solver`` "auto" (direct FFT solve of a fully wet rectangle, direct
                     capacitance solve of a mask inside one, multigrid CG
                     otherwise), "mg", "fft" or "cap"
* ``poisson_warm``   FP32 direct solves refine from the solution held in x
                     (the previous stage's streamfunction)
* ``poisson_refine`` "last" (FP32): only the last RK stage's direct solve of
                     each step refines in FP64, the others are plain FP32
                     solves -- eulerpsi's FFT solves (1.7e-6 from FP64
                     after 1000 steps at 1024^2, against 5.1e-7) and
                     qgrsw's capacitance solves; "every": every stage
                     refines.  Host-checked solves always refine
* ``poisson_opts``   dict of multigrid schedule options, as picked by
                     ``fluids2d_amd.autotune``
* ``tile``           LDS tile code of the fused transport stage (0 16x64,
                     1 8x64, 2 16x32, 3 32x32, 4 8x128); results do not
                     depend on it
* ``hipgraph``       replay runs of device steps from captured HIP graphs
* ``device``         torch device index of the GPU
"""
import copy

MODELS = ("euler", "eulerpsi", "advection", "vectoradv", "boussinesq", "hydrostatic",
          "rsw", "qgrsw", "qg")
FLUX_METHODS = ("weno", "upwind", "centered", "cweno")
INNER_METHODS = FLUX_METHODS + ("classic",)
INTEGRATORS = ("rk3", "ef", "enrk3", "LFRA")
PRECISIONS = ("fp64", "fp32")
SOLVERS = ("auto", "mg", "fft", "cap")

# models whose rhs/diag run on the GPU path
GPU_MODELS = ("euler", "eulerpsi", "advection", "qg", "rsw", "qgrsw")

#          name            default       accepted       help
SCHEMA = (
    ("model",          "euler",      MODELS,        "equations solved"),
    ("nx",             40,           None,          "cells in x"),
    ("ny",             40,           None,          "cells in y"),
    ("Lx",             1.0,          None,          "domain length in x"),
    ("Ly",             1.0,          None,          "domain length in y"),
    ("xperiodic",      False,        None,          "periodic in x"),
    ("yperiodic",      False,        None,          "periodic in y"),
    ("halowidth",      3,            None,          "halo cells on each side"),
    ("noslip",         None,         None,          "None, True or a list of walls"),
    ("f0",             10.0,         None,          "Coriolis parameter"),
    ("beta",           0.0,          None,          "planetary vorticity gradient"),
    ("g",              1,            None,          "gravity"),
    ("H",              1,            None,          "layer depth"),
    ("tend",           1.0,          None,          "end time"),
    ("dt",             0.0,          None,          "fixed step (0: CFL step)"),
    ("maxite",         100,          None,          "iteration cap"),
    ("dtmax",          9e99,         None,          "largest CFL step"),
    ("nprint",         1,            None,          "iterations between prints"),
    ("nplot",          5,            None,          "iterations between frames"),
    ("animation",      False,        None,          "live plot"),
    ("generate_mp4",   False,        None,          "movie of the frames"),
    ("plotvar",        None,         None,          "variable plotted"),
    ("clims",          None,         None,          "colour limits"),
    ("cmap",           "RdBu_r",     None,          "colour map"),
    ("outputfile",     "history.nc", None,          "NetCDF history file"),
    ("var_to_store",   [],           None,          "variables in the history"),
    ("nhis",           0,            None,          "iterations between records"),
    ("integrator",     "rk3",        INTEGRATORS,   "time scheme"),
    ("cfl",            0.9,          None,          "CFL number"),
    ("RAgamma",        0.1,          None,          "Robert-Asselin filter"),
    ("compflux",       "weno",       FLUX_METHODS,  "U*q"),
    ("vortexforce",    "weno",       FLUX_METHODS,  "omega x U"),
    ("innerproduct",   "weno",       INNER_METHODS, "U.u"),
    ("maxorder",       6,            None,          "largest stencil order"),
    ("tracer",         None,         None,          "name of a passive tracer"),
    ("nthreads",       1,            None,          "host threads of the reference"),
    # the MI355X path
    ("precision",      "fp64",       PRECISIONS,    "arithmetic of the device path"),
    ("poisson_rtol",   None,         None,          "Poisson relative residual"),
    ("poisson_maxit",  None,         None,          "Poisson iteration cap"),
    ("poisson_nu",     2,            None,          "multigrid smoothing sweeps"),
    ("poisson_solver", "auto",       SOLVERS,       "Poisson method"),
    ("poisson_warm",   True,         None,          "FP32 direct solves from x"),
    ("poisson_refine", "last",       ("every", "last"),
                                                    "FP64 refinement of FP32 solves"),
    ("poisson_opts",   None,         None,          "multigrid schedule options"),
    ("tile",           0,            None,          "transport LDS tile code"),
    ("hipgraph",       True,         None,          "replay captured step graphs"),
    ("device",         0,            None,          "GPU index"),
)

QUIET = False  # suppress the help text printed by the constructor

_NAMES = tuple(row[0] for row in SCHEMA)


class Param:
    """All parameters as plain attributes, set to their defaults."""

    def __init__(self):
        for name, default, _, _ in SCHEMA:
            setattr(self, name, copy.copy(default))
        if not QUIET:
            self.help()

    def parameters(self):
        """The declared names: the table's, then those of add_parameter."""
        return list(_NAMES) + list(self.__dict__.get("_added", ()))

    def add_parameter(self, name):
        """Declare one more parameter (value None)."""
        setattr(self, name, None)
        if name not in _NAMES:
            added = self.__dict__.setdefault("_added", [])
            if name not in added:
                added.append(name)

    def check_parameters_are_known(self):
        declared = set(self.parameters()) | {"_added"}
        unknown = sorted(n for n in vars(self) if n not in declared)
        assert not unknown, (f"unknown parameter(s) {', '.join(unknown)}; "
                             f"declare them with add_parameter()")
        return True

    def check(self):
        for name, _, accepted, _ in SCHEMA:
            if accepted is not None:
                value = getattr(self, name)
                assert value in accepted, (f"{name} = {value!r} is not one of "
                                           f"{', '.join(accepted)}")
        self.check_parameters_are_known()

    def help(self):
        lines = ["Valid values for string parameters"]
        for name, _, accepted, what in SCHEMA:
            if accepted is not None:
                lines.append(f"  - {name} ({what}): " + ", ".join(accepted))
        print("\n".join(lines) + "\n")
