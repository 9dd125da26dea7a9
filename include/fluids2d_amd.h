/*
 * fluids2d_amd.h -- C ABI of the MI355X (gfx950) time-stepping hot path of
 * Fluids2d.  Implemented in the .hip files of fluids2d_amd/csrc, built into
 * fluids2d_amd/libfluids2d_amd.so.
 *
 * Conventions (all entry points):
 *   - Return 0 on success, a negative code on error; f2d_last_error() gives
 *     the message of the last error on the calling thread.
 *   - Pointers named in the signatures are device pointers unless the name
 *     ends in `_host`.  Arrays are row-major, C-contiguous, of the padded
 *     shape (ny + 2*halowidth, nx + 2*halowidth) unless stated otherwise.
 *   - `dtype` is F2D_F64 or F2D_F32 and applies to every floating array of
 *     the call.  Masks and stencil orders are int8, index maps int32.
 *   - `stream` is a hipStream_t (NULL = the null stream).  No entry point
 *     synchronises except where stated; none allocates on the stepping path
 *     (the Poisson handle allocates its workspace in f2d_poisson_create).
 *   - The caller owns every array; outputs are written in place.
 */
#ifndef FLUIDS2D_AMD_H
#define FLUIDS2D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define F2D_F64 0
#define F2D_F32 1

/* flux families, weno.py:338-343 */
#define F2D_WENO 0
#define F2D_UPWIND 1
#define F2D_CENTERED 2
#define F2D_CWENO 3

/* integrators, integrators.py:127-131 */
#define F2D_RK3 0
#define F2D_EF 1
#define F2D_ENRK3 2

/* Poisson precision modes */
#define F2D_POISSON_F64 0        /* FP64 MG-PCG                                 */
#define F2D_POISSON_F32_REFINE 1 /* FP32 MG-PCG inside FP64 iterative refinement */
#define F2D_POISSON_F32 2        /* FP32 MG-PCG only                             */

/* ABI version of this header (no counterpart in the reference). */
int f2d_abi_version(void);
/* Message of the last failed call on this thread.  Every entry point returns
 * an int status (0 = success) where the reference raises a Python exception
 * (no counterpart in the reference: SuperLU's RuntimeError, elliptic.py:78,
 * and numba's TypeError are its only error paths). */
const char* f2d_last_error(void);
/* Synchronise the current device (host-side wait; no counterpart in the
 * reference, whose kernels are synchronous). */
int f2d_device_synchronize(void);

/* ------------------------------------------------------------------------
 * Grid description shared by the stencil entry points.
 * ------------------------------------------------------------------------ */
typedef struct f2d_grid {
  int64_t ny, nx;   /* interior cells                                   */
  int32_t halowidth;/* must be 3 (param.py:23)                          */
  int32_t xperiodic, yperiodic;
  double dx, dy;    /* Lx/nx, Ly/ny (meshes.py:24-25)                    */
  int32_t tile;     /* LDS tile of the fused transport stage (no
                       counterpart in the reference): 0 16x64 (default),
                       1 8x64, 2 16x32, 3 32x32, 4 8x128 rows x columns;
                       results do not depend on it                      */
} f2d_grid;

/* ------------------------------------------------------------------------
 * Mesh: masks, stencil orders, slip coefficient, unknown numbering.
 * ------------------------------------------------------------------------ */
/* Mesh.set_masks, meshes.py:77-86: mskx, msky, mskv from msk (padded). */
int f2d_set_masks(const int8_t* msk, int64_t ny_pad, int64_t nx_pad,
                  int8_t* mskx, int8_t* msky, int8_t* mskv, void* stream);
/* set_order, meshes.py:146-186, on the flattened array of n entries. */
int f2d_set_order(const int8_t* m, int64_t n, int64_t shift, int maxorder,
                  int8_t* o, void* stream);
/* get_slipcoef, noslip.py:4-36.  mode 0 free-slip, 1 no-slip everywhere,
 * 2 free-slip with no-slip on the sides flagged in `sides` (bit 0 left,
 * 1 right, 2 bottom, 3 top).  coef is float64 of the padded shape. */
int f2d_slipcoef(const int8_t* msk, int64_t ny_pad, int64_t nx_pad, int mode,
                 int sides, int halowidth, double* coef, void* stream);
/* get_Gindex, elliptic.py:198-203: G = running index of wet points,
 * -1 elsewhere.  *count_host receives the number of wet points
 * (synchronises `stream`). */
int f2d_gindex(const int8_t* m, int64_t n, int32_t* G, int64_t* count_host,
               void* stream);

/* ------------------------------------------------------------------------
 * Flat-index WENO loops (weno.py:412-436), one thread per padded point.
 * ------------------------------------------------------------------------ */
/* weno.compflux(flx, U, q, o, s, method), weno.py:412-420 */
int f2d_compflux(int dtype, void* flx, const void* U, const void* q,
                 const int8_t* o, int64_t n, int64_t s, int method,
                 void* stream);
/* weno.vortexforce(du, V, q, o, s, s2, sign, method), weno.py:423-431 */
int f2d_vortexforce(int dtype, void* du, const void* V, const void* q,
                    const int8_t* o, int64_t n, int64_t s, int64_t s2,
                    int sign, int method, void* stream);
/* weno.innerproduct(ke, U, q, o, s, method), weno.py:434-436 (ke += ...) */
int f2d_innerproduct(int dtype, void* ke, const void* U, const void* q,
                     const int8_t* o, int64_t n, int64_t s, int method,
                     void* stream);

/* ------------------------------------------------------------------------
 * Fused transport stage: divflux + fill + RK addto in one pass.
 *   operators.py:21-29 (divflux), :104-107 (div), meshes.py:106-111 (fill),
 *   integrators.py:92-124, 154-196 (rk3/ef/enrk3 stage `stage`, addto).
 * Computes ds = fill(-div(flux(q, U)) * msk) and
 *   q_out = q_in + ((c0*ds0 + c1*ds1) + c2*ds)
 * with the stage coefficients of `integrator` evaluated from *dt (device
 * scalar, float64) exactly as the reference's Python expressions.
 * ds_prev[k] are the tendencies of the earlier stages; ds_out (may be NULL)
 * receives this stage's tendency.  q_out must not alias q_in.
 * With integrator < 0 no update is made (q_out unused): rhs only.
 * ------------------------------------------------------------------------ */
int f2d_transport_stage(int dtype, const f2d_grid* g, int method,
                        const void* q_in, void* q_out, const void* Ux,
                        const void* Uy, const int8_t* ocx, const int8_t* ocy,
                        const int8_t* msk, const void* ds_prev0,
                        const void* ds_prev1, void* ds_out, int integrator,
                        int stage, const double* dt, void* stream);

/* f2d_transport_stage that also copies *dt to *dt_log (one double).
 * This has no counterpart in the reference: it lets a step record its time
 * step without a launch of its own (a captured step's dt log). */
int f2d_transport_stage_log(int dtype, const f2d_grid* g, int method, const void* q_in,
                            void* q_out, const void* Ux, const void* Uy, const int8_t* ocx,
                            const int8_t* ocy, const int8_t* msk, const void* ds_prev0,
                            const void* ds_prev1, void* ds_out, int integrator, int stage,
                            const double* dt, double* dt_log, void* stream);

/* The same stage with U = perpgrad(psi) formed on the fly rather than read:
 * perpgrad, operators.py:144-149 (Ux = -(psi[j+1] - psi[j]) mskx sx,
 * Uy = (psi[i+1] - psi[i]) msky sy, with fill != 0 the periodic halo fill
 * of meshes.py:135-143), bit for bit what f2d_perpgrad_ex would store.
 * A stage whose velocity is not needed elsewhere then skips writing U. */
int f2d_transport_stage_psi(int dtype, const f2d_grid* g, int method, const void* q_in,
                            void* q_out, const void* psi, const int8_t* mskx,
                            const int8_t* msky, double sx, double sy, int fill,
                            const int8_t* ocx, const int8_t* ocy, const int8_t* msk,
                            const void* ds_prev0, const void* ds_prev1, void* ds_out,
                            int integrator, int stage, const double* dt, void* stream);

/* The same stage with the contravariant velocity formed from the covariant
 * one: U = sharp(u), operators.py:59-64 (Ux = ux / dx**2, Uy = uy / dy**2
 * as k_sharp stores them), bit for bit the U that f2d_sharp would store.
 * The shallow-water stage uses it so that U need not be stored. */
int f2d_transport_stage_u(int dtype, const f2d_grid* g, int method, const void* q_in,
                          void* q_out, const void* ux, const void* uy, const int8_t* ocx,
                          const int8_t* ocy, const int8_t* msk, const void* ds_prev0,
                          const void* ds_prev1, void* ds_out, int integrator, int stage,
                          const double* dt, void* stream);

/* f2d_transport_stage_log whose dt is the CFL step of model.py:84-87 formed
 * from the block maxima `part` that f2d_perpgrad_part left (npart entries,
 * f2d_cfl_partials of this grid) with `cfl` and `dtmax`: every workgroup
 * forms it (the maxima are exact, so all get the same dt, bit for bit the
 * one f2d_cfl_finish would write), uses it for this stage, and stores it
 * in *dt for the step's later stages (and in *dt_log if not NULL).  The
 * step loop of model.py:52-69 then needs no launch of its own for the dt:
 * the last stage's perpgrad leaves the maxima, the next step's first stage
 * finishes them. */
int f2d_transport_stage_cfl(int dtype, const f2d_grid* g, int method, const void* q_in,
                            void* q_out, const void* Ux, const void* Uy, const int8_t* ocx,
                            const int8_t* ocy, const int8_t* msk, const void* ds_prev0,
                            const void* ds_prev1, void* ds_out, int integrator, int stage,
                            double* dt, const double* part, int64_t npart, double cfl,
                            double dtmax, double* dt_log, void* stream);

/* addto, integrators.py:154-196 (the sum of :174): over n entries
 * y += ((c0*x0 + c1*x1) + c2*x2), with the coefficients of `integrator`
 * stage `stage` evaluated from *dt as above. */
int f2d_rk_addto(int dtype, void* y, const void* x0, const void* x1,
                 const void* x2, int64_t n, int integrator, int stage,
                 const double* dt, void* stream);
/* addto, integrators.py:154-196, with given coefficients (the LFRA
 * integrator, integrators.py:38-53): over n entries y += ((c0 x0 + c1 x1) +
 * c2 x2) for the first nterms terms, coefficient i multiplied by *dt (device
 * float64) when bit i of dtmask is set. */
int f2d_addto(int dtype, void* y, const void* x0, const void* x1, const void* x2,
              int64_t n, int nterms, double c0, double c1, double c2, const double* dt,
              int dtmask, void* stream);

/* centerstovertices, operators.py:126-133 (addto=0/1) */
int f2d_c2v(int dtype, const f2d_grid* g, const void* c, void* v,
            const int8_t* mskv, int addto, void* stream);
/* perpgrad (+ fill), operators.py:144-149, equations.py:79-84.
 * If maxabs (device float64[2]) is not NULL it receives
 * (max|Ux|, max|Uy|) over the padded arrays. */
int f2d_perpgrad(int dtype, const f2d_grid* g, const void* psi, void* Ux,
                 void* Uy, const int8_t* mskx, const int8_t* msky,
                 int contravariant, double* maxabs, void* stream);
/* perpgrad with the scale factors given by the caller (the reference's
 * 1/mesh.dy**2 and 1/mesh.dx**2 as Python evaluates them, or 1.0 for the
 * covariant form); fill != 0 applies the halo fill; dt (device float64, may
 * be NULL) receives the CFL step of model.py:84-87 computed from the new
 * velocities with `cfl` and `dtmax`. */
int f2d_perpgrad_ex(int dtype, const f2d_grid* g, const void* psi, void* Ux,
                    void* Uy, const int8_t* mskx, const int8_t* msky,
                    double sx, double sy, int fill, double* maxabs,
                    double* dt, double cfl, double dtmax, void* stream);
/* f2d_perpgrad_ex with the CFL step or the maxima requested (maxabs or dt
 * not NULL), allocating nothing: no counterpart in the reference.  It is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops).  The block maxima go to a fixed 128 KiB array
 * in each GPU's memory, held until the process ends, instead of the
 * reduction scratch the library allocates on first use.  Limit: at most
 * 8192 x 1024 = 8388608 padded points, (ny+6)(nx+6), about 2890^2 cells.
 * Larger grids are refused with an error.  Calls on one GPU share the
 * array, so calls that overlap on one GPU (concurrent streams) race.  The
 * first call on each GPU looks up the array's device address; make it
 * outside a stream capture.  Results are bit for bit f2d_perpgrad_ex's. */
int f2d_perpgrad_cfl_noalloc(int dtype, const f2d_grid* g, const void* psi, void* Ux,
                             void* Uy, const int8_t* mskx, const int8_t* msky,
                             double sx, double sy, int fill, double* maxabs,
                             double* dt, double cfl, double dtmax, void* stream);
/* The CFL step of Model.set_dt (model.py:84-87) in two parts.
 * f2d_cfl_partials: the number of float64 entries (*npart) of the block
 * maxima of |Ux| and |Uy| that f2d_perpgrad_part writes for grid g.
 * f2d_perpgrad_part: f2d_perpgrad_ex's gradient, with the block maxima
 * left in the caller's device array `part` (npart entries) instead of
 * being reduced.  f2d_cfl_finish: the reduction, dt = min(cfl / (max|Ux| +
 * max|Uy| + 1e-99), dtmax) into *dt and the maxima into maxabs[2] (either
 * may be NULL), bit for bit f2d_perpgrad_ex's; f2d_transport_stage_cfl
 * forms the same dt inside the next stage instead. */
int f2d_cfl_partials(const f2d_grid* g, int64_t* npart);
/* perpgrad, operators.py:144-149, with the block maxima of the CFL step
 * (model.py:84-87) left in part */
int f2d_perpgrad_part(int dtype, const f2d_grid* g, const void* psi, void* Ux, void* Uy,
                      const int8_t* mskx, const int8_t* msky, double sx, double sy, int fill,
                      double* part, int64_t npart, void* stream);
/* the CFL step of model.py:84-87 from those block maxima */
int f2d_cfl_finish(const double* part, int64_t npart, double* maxabs, double* dt, double cfl,
                   double dtmax, void* stream);
/* fill_halo_array, meshes.py:135-143 (+ the y-periodic extension) */
int f2d_fill_halo(int dtype, const f2d_grid* g, void* a, void* stream);
/* Model.set_dt, model.py:71-87 (non-rsw branch):
 * dt = min(cfl / (max|Ux| + max|Uy| + 1e-99), dtmax) over the padded
 * arrays; result in *dt (device float64). */
int f2d_cfl_dt(int dtype, const void* Ux, const void* Uy, int64_t n,
               double cfl, double dtmax, double* dt, void* stream);
/* f2d_cfl_dt (Model.set_dt, model.py:84-87) with its block maxima in the
 * caller's device array `part` (npart >= 2048 entries, or 2 ceil(n / 256)),
 * bit for bit the same dt.  Unlike f2d_cfl_dt it may be recorded into a
 * member batch (f2d_record_begin): each member brings its own part. */
int f2d_cfl_dt_part(int dtype, const void* Ux, const void* Uy, int64_t n, double cfl,
                    double dtmax, double* part, int64_t npart, double* dt, void* stream);
/* dt = value (device scalar): the fixed-dt and rsw branches of Model.set_dt,
 * model.py:71-82 */
int f2d_set_scalar(double* dst, double value, void* stream);
/* Give `stream` (on the current device) its own scratch for the grid
 * reductions of f2d_cfl_dt and f2d_perpgrad_ex, allocated here.  It has
 * no counterpart in the reference: it is for callers that run models on
 * concurrent streams or graph branches.  Without it all streams of a
 * device share one buffer, and reductions on different streams must not
 * overlap. */
int f2d_reserve_stream_scratch(void* stream);
/* dst[0:n] = src[0:n] on the device, as a kernel (a graph-capturable copy
 * of device scalars such as the step's dt into a log): the device form of
 * the per-step dt the host reads in Model.set_dt, model.py:71-87 */
int f2d_copy_f64(double* dst, const double* src, int n, void* stream);

/* ------------------------------------------------------------------------
 * C-grid point operators of the shallow-water models (operators.py), each
 * writing exactly the entries the reference's numpy slices assign (the
 * others keep their values), masks applied by multiplication, sums left to
 * right.  Padded arrays of the grid; dtype F2D_F64 or F2D_F32.
 * ------------------------------------------------------------------------ */
/* addcoriolis, operators.py:32-39: dux += f*(4 Uy), duy -= f*(4 Ux),
 * f = f0*area*0.25 */
int f2d_coriolis(int dtype, const f2d_grid* g, const void* Ux, const void* Uy,
                 void* dux, void* duy, double f, void* stream);
/* addgrad, operators.py:49-56: du -= grad(phi) * (mskx, msky) */
int f2d_addgrad(int dtype, const f2d_grid* g, const void* phi,
                const int8_t* mskx, const int8_t* msky, void* dux, void* duy,
                void* stream);
/* sharp, operators.py:59-64: U = u * (1/dx**2, 1/dy**2) */
int f2d_sharp(int dtype, const f2d_grid* g, const void* ux, const void* uy,
              void* Ux, void* Uy, void* stream);
/* compute_vorticity, operators.py:67-77 (slipcoef: float64 array) */
int f2d_vorticity(int dtype, const f2d_grid* g, const void* ux, const void* uy,
                  const double* slipcoef, void* omega, void* stream);
/* Failure detection (no counterpart in the reference, which has no NaN
 * check; its loop stops only on SIGINT, model.py:33-45): sets *flag
 * (int32, device) to 1 if an entry of the padded array `a` where msk is
 * wet (msk NULL: any entry) is not finite.  Never clears the flag. */
int f2d_nonfinite(int dtype, const f2d_grid* g, const void* a, const int8_t* msk,
                  int32_t* flag, void* stream);
/* the last statement of compute_kinetic_energy, operators.py:99-101:
 * ke *= msk * 0.5 (the inner products are f2d_innerproduct) */
int f2d_ke_finish(int dtype, const f2d_grid* g, void* ke, const int8_t* msk,
                  void* stream);
/* compute_pressure, operators.py:110-111: p = (g/area) * (h + hb) */
int f2d_pressure(int dtype, const f2d_grid* g, const void* h, void* p,
                 double g_over_area, double hb, void* stream);
/* div, operators.py:104-107: delta = -(div F) * msk (last row/column of
 * the x/y parts as numpy leaves them) */
int f2d_div(int dtype, const f2d_grid* g, const void* Fx, const void* Fy,
            const int8_t* msk, void* delta, void* stream);
/* centerstovertices of c*scale, operators.py:126-133 (addto as the
 * reference's flag), as add_stretching calls it, operators.py:194-201 */
int f2d_c2v_scaled(int dtype, const f2d_grid* g, const void* c, double scale,
                   void* v, const int8_t* mskv, int addto, void* stream);
/* verticestocenters of vh*scale, operators.py:136-141, as
 * thickness_from_psi calls it, operators.py:204-211 (cells with no wet
 * vertex get (1/0)*0 = NaN before the mask, as in numpy) */
int f2d_v2c_scaled(int dtype, const f2d_grid* g, const void* vh, double scale,
                   const int8_t* mskv, const int8_t* msk, void* h, void* stream);

/* ------------------------------------------------------------------------
 * Fused shallow-water stage (rsw, equations.py:139-157; qgrsw,
 * equations.py:106-136): the operator sequence of a stage in few launches,
 * each entry computed with the arithmetic and order of the one-operator
 * entry points above, so FP64 results are bit for bit theirs.  Fields of
 * a stage: 0 u.x, 1 u.y, 2 h.
 * ------------------------------------------------------------------------ */
typedef struct f2d_sw_stage {
  int32_t integrator; /* F2D_RK3/F2D_EF/F2D_ENRK3; -1: tendency only       */
  int32_t stage;      /* stage index (integrators.py:92-124)               */
  const double* dt;   /* device float64 scalar: the step's dt              */
  const void* prev[2][3]; /* tendencies of the earlier stages, [stage][field] */
  void* out[3];       /* this stage's tendency (NULL: not stored)          */
  void* y[3];         /* fields updated in place, y += the RK combination
                         of addto, integrators.py:154-196 (NULL: none)     */
  const void* y_in[3]; /* non-NULL: y = y_in + the combination instead
                         (y_in left unchanged: a ping-pong pair)           */
  double* dt_log;     /* f2d_sw_momentum copies *dt here (NULL: no copy;
                         dt must be given, also with integrator -1)        */
} f2d_sw_stage;

/* The momentum tendency: vortexforce of both components (weno.py:423-431,
 * V = U.y resp. U.x, orders ovx/ovy), addcoriolis with f = f0*area*0.25
 * (operators.py:32-39), addgrad(ke) then addgrad(p) when ke and p are
 * given (operators.py:49-56), the halo fill when fill != 0 (meshes.py:
 * 135-143), then the stage terms of fields 0 and 1 (st may be NULL: the
 * tendency goes nowhere).  Reads no u: st->y may be u itself. */
int f2d_sw_momentum(int dtype, const f2d_grid* g, int method, const void* Ux, const void* Uy,
                    const void* omega, const int8_t* ovx, const int8_t* ovy, double f,
                    const void* ke, const void* p, const int8_t* mskx, const int8_t* msky,
                    int fill, const f2d_sw_stage* st, void* stream);
/* f2d_sw_momentum with U = sharp(u) (operators.py:59-64) and p =
 * compute_pressure(h) (operators.py:110-111, p = (g/area)(h + hb)) formed
 * from u and h on the fly, with the arithmetic of f2d_sharp and
 * f2d_pressure: bit for bit the same tendency, without U and p stored.
 * u is read around each point, so the update of fields 0 and 1 goes to
 * other arrays (st->y != u, st->y_in = u). */
int f2d_sw_momentum_u(int dtype, const f2d_grid* g, int method, const void* ux,
                      const void* uy, const void* omega, const int8_t* ovx, const int8_t* ovy,
                      double f, const void* ke, const void* h, double g_over_area, double hb,
                      const int8_t* mskx, const int8_t* msky, int fill,
                      const f2d_sw_stage* st, void* stream);
/* The diag of rsw (equations.py:151-157) and qgrsw (equations.py:130-136):
 * sharp (operators.py:59-64), compute_vorticity (operators.py:67-77),
 * compute_kinetic_energy with `method` (operators.py:80-101) when ke is
 * given, compute_pressure (operators.py:110-111) when p is given, and the
 * fill of omega and ke.  Ux and Uy may be NULL (sharp not stored). */
int f2d_sw_diag(int dtype, const f2d_grid* g, int method, const void* ux, const void* uy,
                const void* h, const double* slipcoef, const int8_t* okx, const int8_t* oky,
                const int8_t* msk, double g_over_area, double hb, void* Ux, void* Uy,
                void* omega, void* ke, void* p, void* stream);
/* qg_projection's right-hand side (operators.py:176-180, 194-201) from the
 * unfilled, unprojected tendencies: pv = vorticity(du) + c2v(dh*scale),
 * times mskv. */
int f2d_qg_pv(int dtype, const f2d_grid* g, const void* dux, const void* duy, const void* dh,
              const double* slipcoef, const int8_t* mskv, double scale, void* pv,
              void* stream);
/* The rest of qg_projection after the Helmholtz solve (operators.py:181-
 * 183, 204-211) and of the stage: dh = v2c(psi*scale)*msk, du =
 * perpgrad(psi) (covariant), the entries the slices leave keeping the
 * unprojected tendency pre_*, fill, then the stage terms of the three
 * fields.  st->out and st->y must not be the pre_* arrays. */
int f2d_qg_project(int dtype, const f2d_grid* g, const void* psi, const int8_t* mskx,
                   const int8_t* msky, const int8_t* mskv, const int8_t* msk, double scale,
                   const void* pre_ux, const void* pre_uy, const void* pre_h,
                   const f2d_sw_stage* st, void* stream);

/* Bulk diagnostics, diagnostics.py:45-66 (Bulk.__call__): out[0..3] =
 * ke, ens, vort, angular from FP64 sums over every point of the padded
 * arrays (numpy's domsum), on the device; ke NULL (eulerpsi, whose state
 * has no ke) takes the cell-centred 1/2 u.U of the contravariant U.
 * partials: scratch of at least 6 * nparts doubles (nparts blocks at most). */
int f2d_bulk_diagnostics(int dtype, const f2d_grid* g, const void* ke, const void* omega,
                         const void* Ux, const void* Uy, const int8_t* msk, double* partials,
                         int nparts, double* out, void* stream);

/* ------------------------------------------------------------------------
 * Adjoint of the eulerpsi RK stage on a fully wet, doubly periodic grid (no
 * counterpart in the reference, which has no derivatives): the stage
 * evaluates R(s) = -div(U r(s)) with U = perpgrad(A^-1 c2v(s)) (weno.py:346-
 * 364, operators.py:126-149, elliptic.py:80-87); these entry points build
 * R'(s)^T lam in two calls around the solve: f2d_adj_stage (transport and
 * perpgrad reversed), the Poisson solve of lpsi (A^-1 is self-adjoint), then
 * f2d_adj_combine (c2v reversed).  Adjoint arrays are read on the interior
 * (periodic neighbours wrapped) and written on the interior of padded
 * arrays; the forward state q, Ux, Uy need their periodic halos filled.
 * ------------------------------------------------------------------------ */
/* Transport part (compflux/div of weno.py:346-364, operators.py:21-29,
 * reversed) and perpgrad^T (operators.py:144-149 reversed, contravariant
 * factors) for the stage adjoint lam = c0 lam0 + c1 lam1 + c2 lam2 (lam1,
 * lam2 may be NULL), the RK combination of integrators.py:92-107 reversed:
 * lq = (dT/dq)^T lam, lUx/lUy = (dT/dU)^T lam, lpsi = perpgrad^T (lUx, lUy).
 * work: 12 padded arrays of scratch. */
int f2d_adj_stage(int dtype, const f2d_grid* g, const void* q, const void* Ux,
                  const void* Uy, const void* lam0, const void* lam1, const void* lam2,
                  double c0, double c1, double c2, void* lUx, void* lUy, void* work,
                  void* lq, void* lpsi, void* stream);
/* out = base + lq + c2v^T lb, c2v of operators.py:126-133 reversed.  out
 * may be base or any other array (interior written). */
int f2d_adj_combine(int dtype, const f2d_grid* g, const void* base, const void* lq,
                    const void* lb, void* out, void* stream);

/* ------------------------------------------------------------------------
 * Masked Poisson / Helmholtz on the vertex grid (elliptic.py:71-203).
 * Solves A x = b at the wet points of `unknowns_host` where A is the
 * masked 5-point operator of get_Laplacian_sparse ("v", Dirichlet):
 *   (A x)_P = (dy/dx)(x_E + x_W) + (dx/dy)(x_N + x_S)
 *             - (2(dy/dx + dx/dy) + maindiag) x_P,
 * missing (dry) neighbours contributing zero, with x-periodic (and, as an
 * extension, y-periodic) wrap.  The doubly periodic operator without mass
 * term is singular: b is projected to zero mean and x has zero mean.
 * Geometric multigrid (red-black Gauss-Seidel V-cycle) preconditioned CG.
 * ------------------------------------------------------------------------ */
typedef struct f2d_poisson f2d_poisson;

typedef struct f2d_poisson_opts {
  int32_t mode;          /* F2D_POISSON_*                                 */
  int32_t nu;            /* RB-GS sweeps before and after (1 or 2)        */
  int32_t coarse_sweeps; /* symmetric sweeps on the coarsest level (0=auto) */
  int32_t refine_inner;  /* F32_REFINE: FP32 PCG iterations per refinement */
  int32_t tail_points;   /* max points of the single-workgroup tail levels
                            (0: 1024, levels from 32^2 down)             */
  int32_t tail_threads;  /* workgroup size of the tail kernel (256/512/1024) */
  int32_t nu_coarse;     /* sweeps on the multi-block coarse levels (0: nu) */
  int32_t nu_tail;       /* sweeps on the tail levels (0: nu)              */
  int32_t solver;        /* F2D_SOLVER_*: 0 auto, 1 multigrid CG, 2 FFT,
                            4 capacitance                                  */
  int32_t fft_refine;    /* FFT in FP32 modes: FP64 refinements (0: 1)     */
  int32_t fft_warm;      /* FFT in F32_REFINE: 1 = from the handle's second
                            solve on, the first refinement starts from x
                            (its interior) instead of a cold FP32 solve;
                            0 = cold (default)                             */
  int32_t neumann;       /* 1: the Neumann ("c") operator of
                            get_Laplacian_sparse (elliptic.py:186-190,
                            diagonal minus the sum of the unknown
                            neighbours' coefficients minus maindiag).  Every
                            cell wet and power-of-two sizes: the direct
                            solve (even extensions about closed walls; zero
                            mean when singular) unless solver is MG; else
                            Jacobi-preconditioned CG in FP64 (x keeps the
                            constant of its initial guess)                 */
} f2d_poisson_opts;

/* Solver selection.  AUTO takes the direct spectral (FFT) solve when every
 * vertex strictly inside the domain is an unknown and nx, ny are powers of
 * two in [8, 4096] ([8, 2048] in a closed direction, solved as the odd
 * extension on the doubled periodic grid; the operator is then diagonalised
 * by the 2-D DFT) and, in F2D_POISSON_F64, the transformed (doubled when
 * closed) sizes are at most 2048 (one row in a workgroup's LDS); the
 * multigrid-preconditioned CG otherwise. */
#define F2D_SOLVER_AUTO 0
#define F2D_SOLVER_MG 1
#define F2D_SOLVER_FFT 2
#define F2D_SOLVER_NEUMANN_CG 3  /* reported by f2d_poisson_method only */
/* The masked vertex operator of a mask inside a rectangle the FFT solve
 * handles (power-of-two sizes, a closed direction or maindiag > 0): the
 * masked operator is the rectangle operator's principal submatrix on the
 * unknowns, so x = A_R^-1 (b + s) with a correction s on the dry points
 * next to the unknowns from the inverse of their capacitance matrix (formed
 * once at creation with rocSOLVER).  A direct solve: two FFT solves and one
 * dense matrix-vector product.  "auto" takes it for at most 8192 such points
 * (FP64 and F32_REFINE modes). */
#define F2D_SOLVER_CAP 4

/* unknowns_host: int8, padded shape, 1 at unknowns (elliptic.py:102-111:
 * mskv with the periodic halos cleared).  opts may be NULL (defaults). */
int f2d_poisson_create(const int8_t* unknowns_host, const f2d_grid* g,
                       double maindiag, const f2d_poisson_opts* opts,
                       f2d_poisson** out);
/* Frees the handle: the reference's factor lives as long as its Poisson2D
 * object (elliptic.py:71-78). */
int f2d_poisson_destroy(f2d_poisson* h);
/* Poisson2D.solve(b, x), elliptic.py:80-87: x (in/out, initial guess) is
 * written at the unknowns only, then its halo is filled.  The solve stops
 * when ||r||_2 <= rtol ||b||_2 or after maxit iterations.  iters/relres
 * (host, may be NULL) receive the iteration count and the final relative
 * residual; giving either synchronises `stream`. */
int f2d_poisson_solve(f2d_poisson* h, const void* b, void* x, double rtol,
                      int maxit, int* iters, double* relres, void* stream);
/* The method a handle solves with, F2D_SOLVER_MG, F2D_SOLVER_FFT,
 * F2D_SOLVER_NEUMANN_CG or F2D_SOLVER_CAP (-1 for NULL); no counterpart in
 * the reference, whose SuperLU factor is its only method (elliptic.py:78).
 * With FFT or CAP the solve is direct: `maxit` is unused, *iters receives
 * 1 + the FP64 refinements, *relres the measured true relative residual,
 * and the solve counts as converged when that residual is below 1e-6
 * (`rtol` is not its acceptance test: a direct FP64 solve reaches eps times
 * the condition number, about 1e-11 at 1024^2). */
int f2d_poisson_method(f2d_poisson* h);
/* FP64 refinement passes of the handle's following direct (FFT or
 * capacitance) solves in the FP32 modes: 0 makes them plain FP32 solves, -1
 * returns to f2d_poisson_opts.fft_refine (no counterpart in the reference,
 * whose solves are FP64; multigrid ignores it).  A stepper may refine only
 * the last RK stage's solve of each step. */
int f2d_poisson_set_refine(f2d_poisson* h, int nref);
/* Fused eulerpsi diagnostic solve (equations.py:80-82): b = c2v(omega)*mskv
 * computed on the fly from the cell field omega, then solve as above. */
int f2d_poisson_solve_c2v(f2d_poisson* h, const void* omega, void* x,
                          double rtol, int maxit, int* iters, double* relres,
                          void* stream);
/* Device-side solve statistics (float64 device array of 8): of the last
 * solve [0] iterations (F32_REFINE: inner iterations summed), [1] relative
 * residual, [2] converged flag, [3] refinements; since the last reset
 * [4] number of solves that did not converge, [5] maximum iterations
 * (F32_REFINE: refinements), [6] number of solves.  Without iters/relres a
 * solve enqueues exactly `maxit` iterations (refinements), those after
 * convergence returning at once, and never synchronises: these counters are
 * how a caller checks such solves afterwards. */
/* (no counterpart in the reference: SuperLU is direct, elliptic.py:78-87) */
const double* f2d_poisson_stats(f2d_poisson* h);
/* Zero the counters [4..6] (no counterpart in the reference). */
int f2d_poisson_reset_stats(f2d_poisson* h, void* stream);
/* Copy the 8 statistics above to host memory, synchronising `stream`
 * (no counterpart in the reference).  FFT solves made without iters/relres
 * measure nothing: they only add to the solve count [6]. */
int f2d_poisson_read_stats(f2d_poisson* h, double* stats_host, void* stream);
/* Measurement hook (no counterpart in the reference): launch one fine-level
 * kernel of the CG iteration `reps` times on the handle's current vectors --
 * kind 0 the update/pre-smooth/restrict pass, 1 the prolongate/post-smooth
 * pass, 2 the search-direction/operator pass -- with a scratch control block
 * and scratch outputs, not touching the caller's solution.  Kinds 3..9
 * time the passes of the FFT solve instead (3 rows forward, 4 rows forward
 * forming the FP64 refinement residual, 5 columns, 6 rows inverse, 7 rows
 * inverse accumulating the refinement, 8 rows forward of the warm start's
 * residual of an FP32 guess, 9 rows inverse accumulating onto the FP32
 * guess; 4, 7, 8 and 9 in FP32 modes only).  *us receives
 * the mean launch time from HIP events on `stream`, *bytes the algorithmic
 * bytes of one launch (each array read once, written once).  Synchronises. */
int f2d_poisson_bench_kernel(f2d_poisson* h, int kind, int reps, double* us,
                             double* bytes, void* stream);

/* Member-batched solves for ensembles of independent models (no counterpart
 * in the reference, which runs one Poisson2D.solve per model and call,
 * elliptic.py:80-87): n handles of the same rectangle operator (grid,
 * spacings, maindiag, periodicity; masks may differ), each solving with
 * F2D_SOLVER_CAP in an FP32 mode with the same refinement options, bound to
 * fixed device arrays b[i] (right-hand side at the vertices) and x[i]
 * (solution, in/out).  f2d_poisson_batch_solve then performs
 * f2d_poisson_solve(h[i], b[i], x[i], ..., NULL, NULL, stream) for every
 * member -- the same kernels' arithmetic, so the same bits -- with one
 * launch per kernel of the solve for all members.  Nothing synchronises;
 * the solves count in each handle's statistics as solves without
 * iters/relres.  The batch does not own the handles; destroy it first. */
typedef struct f2d_poisson_batch f2d_poisson_batch;
int f2d_poisson_batch_create(f2d_poisson* const* h, int32_t n, const void* const* b,
                             void* const* x, f2d_poisson_batch** out);
/* The solves of every member on `stream` (no counterpart in the reference:
 * Poisson2D.solve per model, elliptic.py:80-87), with `nref` FP64
 * refinements (0: plain FP32 solves; -1: the handles' option). */
int f2d_poisson_batch_solve(f2d_poisson_batch* bt, int nref, void* stream);
/* Frees the batch's device tables (no counterpart in the reference). */
int f2d_poisson_batch_destroy(f2d_poisson_batch* bt);

/* Member batching of kernel launches (no counterpart in the reference, which
 * advances one model per process, model.py:29-69).  Between
 * f2d_record_begin and f2d_record_end the calling thread's calls of
 * f2d_sw_momentum, f2d_sw_momentum_u, f2d_sw_diag, f2d_qg_pv,
 * f2d_qg_project and the f2d_transport_stage family check their arguments
 * and record their kernel launches (kernel, grid and argument values)
 * instead of enqueueing them; every other entry point fails while a
 * recording is open.  f2d_launch_batch_create merges the recordings of n
 * members, which must have recorded the same kernels with the same grids in
 * the same order, into device tables; f2d_launch_batch_run then enqueues
 * one launch per recorded kernel for all members (member i's workgroups
 * run member i's recorded arguments: its launches' results bit for bit,
 * provided no two members write the same arrays).  Recordings and batches
 * are independent of each other once created. */
typedef struct f2d_recording f2d_recording;
typedef struct f2d_launch_batch f2d_launch_batch;
int f2d_record_begin(void);
/* Closes the thread's recording into *out (no counterpart in the
 * reference). */
int f2d_record_end(f2d_recording** out);
/* Number of launches recorded, -1 for NULL (no counterpart in the
 * reference). */
int f2d_recording_size(const f2d_recording* r);
/* Frees a recording (no counterpart in the reference). */
int f2d_recording_destroy(f2d_recording* r);
/* Merges n recordings and synchronises (no counterpart in the reference). */
int f2d_launch_batch_create(f2d_recording* const* recs, int32_t n, f2d_launch_batch** out);
/* Enqueues the batch on `stream` (no counterpart in the reference). */
int f2d_launch_batch_run(f2d_launch_batch* bt, void* stream);
/* Frees the batch's tables, synchronising (no counterpart in the reference). */
int f2d_launch_batch_destroy(f2d_launch_batch* bt);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDS2D_AMD_H */
