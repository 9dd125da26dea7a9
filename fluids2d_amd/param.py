"""Param: flat configuration object, same fields, defaults and validation as
the reference's (param.py:10-109).

Fields added for the MI355X path (declared in the constructor, so that
``check()`` knows them, as the reference's ``add_parameter`` would):

* ``precision``     "fp64" (reproduces the reference) or "fp32" (FP32 stencils
                    and multigrid, FP64 iterative refinement of the Poisson
                    residual)
* ``poisson_rtol``  relative residual of the Poisson solves (None: 1e-13 in
                    FP64, 1e-9 for the refined FP32 mode)
* ``poisson_maxit`` iteration cap of the Poisson solves (None: 100 in FP64,
                    8 refinements in FP32)
* ``poisson_nu``    red-black Gauss-Seidel sweeps before/after each level
* ``poisson_solver`` "auto" (direct FFT solve for the fully wet doubly
                    periodic operator with power-of-two sizes, multigrid CG
                    otherwise), "mg" or "fft"
* ``poisson_warm``  FP32 direct (FFT) solves: the FP64-residual refinement
                    starts from the solution held in x (the previous stage's
                    streamfunction) instead of a cold FP32 solve
* ``poisson_opts``  dict of multigrid schedule options (sweeps per level,
                    tail size and workgroup, inner FP32 iterations), as
                    picked by ``fluids2d_amd.autotune``
* ``tile``          LDS tile code of the fused transport stage (0 16x64,
                    1 8x64, 2 16x32, 3 32x32, 4 8x128), as picked by
                    ``fluids2d_amd.autotune``; results do not depend on it
* ``hipgraph``      replay runs of device steps from captured HIP graphs
* ``device``        torch device index of the GPU
"""
_models = ["euler", "eulerpsi", "advection", "vectoradv",
           "boussinesq", "hydrostatic", "rsw", "qgrsw", "qg"]

_methods = ["weno", "upwind", "centered", "cweno"]
_methods_extended = _methods + ["classic"]

_integrators = ["rk3", "ef", "enrk3", "LFRA"]

# models whose rhs/diag run on the GPU path
_gpu_models = ["euler", "eulerpsi", "advection", "qg", "rsw", "qgrsw"]
_precisions = ["fp64", "fp32"]

QUIET = False  # suppress the help text printed by the constructor


class Param:
    def __init__(self):
        self.model = "euler"

        self.nx = 40
        self.ny = 40

        self.Lx = 1.0
        self.Ly = 1.0

        self.xperiodic = False
        self.yperiodic = False
        self.halowidth = 3
        self.noslip = None

        self.f0 = 10.0
        self.beta = 0.0
        self.g = 1
        self.H = 1

        self.tend = 1.0
        self.dt = 0.0
        self.maxite = 100
        self.dtmax = 9e99
        self.nprint = 1

        self.nplot = 5
        self.animation = False
        self.generate_mp4 = False
        self.plotvar = None
        self.clims = None
        self.cmap = "RdBu_r"

        self.outputfile = "history.nc"
        self.var_to_store = []
        self.nhis = 0

        self.integrator = "rk3"
        self.cfl = 0.9
        self.RAgamma = 0.1

        self.compflux = "weno"
        self.vortexforce = "weno"
        self.innerproduct = "weno"
        self.maxorder = 6

        self.tracer = None

        self.nthreads = 1

        # MI355X path
        self.precision = "fp64"
        self.poisson_rtol = None
        self.poisson_maxit = None
        self.poisson_nu = 2
        self.poisson_solver = "auto"
        self.poisson_warm = True
        self.poisson_opts = None
        self.tile = 0
        self.hipgraph = True
        self.device = 0

        self.__parameters__ = get_parameters(self)
        if not QUIET:
            self.help()

    def add_parameter(self, name):
        setattr(self, name, None)
        self.__parameters__ = get_parameters(self)

    def check_parameters_are_known(self):
        extra = set(get_parameters(self)).difference(set(self.__parameters__))
        if len(extra) == 0:
            return True
        msg = [f"parameter {extra} is unknown",
               f"parameters are {self.__parameters__}"]
        assert False, "\n".join(msg)

    def check(self):
        assert self.model in _models
        assert self.compflux in _methods
        assert self.vortexforce in _methods
        assert self.innerproduct in _methods_extended
        assert self.integrator in _integrators
        assert self.precision in _precisions
        self.check_parameters_are_known()

    def help(self):
        doc = ["Valid values for string parameters"]
        doc += [f"  - {bold('model')}: " + ", ".join(_models)]
        doc += [f"  - {bold('integrator')}: " + ", ".join(_integrators)]
        doc += [f"  - {bold('compflux')} (U*q): " + ", ".join(_methods)]
        doc += [f"  - {bold('vortexforce')} (omega x U): " + ", ".join(_methods)]
        doc += [f"  - {bold('innerproduct')} (U.u): " + ", ".join(_methods_extended)]
        doc += [f"  - {bold('precision')}: " + ", ".join(_precisions)]
        doc += [""]
        print("\n".join(doc))


def bold(s):
    return '\033[1m' + '\033[94m' + s + '\033[0m'


def get_parameters(self):
    return [d for d in self.__dir__() if "__" not in d]
