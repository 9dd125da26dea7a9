// f2d_sw.hip -- fused stage kernels of the shallow-water models (rsw,
// qgrsw, equations.py:106-157): the reference's operator sequence of a
// stage evaluated point by point in few launches, with each operator's
// arithmetic and evaluation order, so FP64 results are bit-identical to the
// one-operator-per-launch path (f2d_ops.hip, f2d_stencil.hip).
//
//   k_sw_momentum   rsw rhs for u: vortexforce (weno.py:367-385), then
//                   addcoriolis, addgrad(ke), addgrad(p) (operators.py:6-56),
//                   fill (meshes.py:135-143), and the RK addto of u
//                   (integrators.py:154-196); qgrsw: vortexforce + coriolis
//                   only, unfilled, into scratch (projected next)
//   k_sw_diag       sharp, compute_vorticity, compute_kinetic_energy,
//                   compute_pressure, fill(omega, ke) (operators.py:59-111)
//   k_qg_pv         qg_projection's vorticity(du) + centerstovertices(dh
//                   (-f0/H)) (operators.py:176-201)
//   k_qg_project    after the Helmholtz solve: thickness_from_psi and
//                   perpgrad (operators.py:144-149, 204-211), fill, addto
//
// A fill gives a periodic halo point its wrapped source's value: the kernel
// computes the source's value there.  Entries numpy's slices leave alone
// keep their old value; where such an entry has periodic images, its own
// thread writes them (an image's thread reading it would race with the
// owner's update).
#include "f2d_common.hpp"

namespace f2d {

// numba's negative-index wrap (and, never reached with valid orders, the
// high side), as the flat loops of f2d_stencil.hip
__device__ __forceinline__ int64_t wrn(int64_t i, int64_t n) {
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

struct SwGeo {
  int NY, NX, ny, nx, xper, yper;
};

__device__ __forceinline__ int wsrc_i(const SwGeo& g, int i) {
  return g.xper ? (i < 3 ? i + g.nx : (i >= g.nx + 3 ? i - g.nx : i)) : i;
}
__device__ __forceinline__ int wsrc_j(const SwGeo& g, int j) {
  return g.yper ? (j < 3 ? j + g.ny : (j >= g.ny + 3 ? j - g.ny : j)) : j;
}

// vortexforce_on_interval at flat index i (weno.py:370-384): V the
// transverse contravariant velocity, q the vertex vorticity, o the order
template <class T>
__device__ __forceinline__ T vforce(int method, const T* V, const T* q, int o, int64_t i,
                                    int64_t s, int64_t s2, T sign, int64_t n) {
  if (o <= 0) return T(0);
  const T Vm = T(0.25) * (((V[i] + V[wrn(i + s, n)]) + V[wrn(i - s2, n)]) +
                          V[wrn(i + s - s2, n)]);
  T f;
  if (o > 4)
    f = f5<T>(method, Vm, q[wrn(i - 2 * s, n)], q[wrn(i - s, n)], q[i], q[wrn(i + s, n)],
              q[wrn(i + 2 * s, n)], q[wrn(i + 3 * s, n)]);
  else if (o > 2)
    f = f3<T>(method, Vm, q[wrn(i - s, n)], q[i], q[wrn(i + s, n)], q[wrn(i + 2 * s, n)]);
  else
    f = f1<T>(method, Vm, q[i], q[wrn(i + s, n)]);
  return (sign * f) * Vm;
}

// innerproduct_on_interval term at flat index i (weno.py:391-403), with the
// contravariant velocity U = u * sc formed from u as sharp stores it
template <class T>
__device__ __forceinline__ T iprod(int method, const T* u, T sc, int o, int64_t i,
                                   int64_t s, int64_t n) {
  const T Um = T(0.5) * (u[i] * sc + u[wrn(i + s, n)] * sc);
  T f;
  if (o > 4)
    f = f5<T>(method, Um, u[wrn(i - 2 * s, n)], u[wrn(i - s, n)], u[i], u[wrn(i + s, n)],
              u[wrn(i + 2 * s, n)], u[wrn(i + 3 * s, n)]);
  else if (o > 2)
    f = f3<T>(method, Um, u[wrn(i - s, n)], u[i], u[wrn(i + s, n)], u[wrn(i + 2 * s, n)]);
  else
    f = f1<T>(method, Um, u[i], u[wrn(i + s, n)]);
  return f * Um;
}

// the RK stage terms of one field: earlier tendencies (up to two), this
// stage's tendency out (nullable), the array updated in place (nullable)
template <class T>
struct Terms {
  const T* p0;
  const T* p1;
  T* out;
  T* y;
};

template <class T>
__device__ __forceinline__ void stage_out(const Terms<T>& t, int nc, const double c[3],
                                          int64_t k, T ds) {
  if (t.out) t.out[k] = ds;
  if (t.y && nc > 0) {
    const T x0 = (nc == 1) ? ds : t.p0[k];
    const T x1 = (nc == 2) ? ds : (nc > 2 ? t.p1[k] : T(0));
    const T x2 = (nc == 3) ? ds : T(0);
    t.y[k] = t.y[k] + rk_combine<T>(nc, c, x0, x1, x2);
  }
}

template <class T>
struct MomArgs {
  SwGeo g;
  int method, fill, integ, stage;
  const T* Ux;
  const T* Uy;
  const T* w;
  const int8_t* ovx;
  const int8_t* ovy;
  double f;  // addcoriolis coefficient f0 * area * 0.25
  const T* ke;  // nullable (qgrsw: no gradient terms)
  const T* p;
  const int8_t* mx;
  const int8_t* my;
  Terms<T> tx, ty;
  const double* dt;
};

template <class T>
__global__ __launch_bounds__(256) void k_sw_momentum(MomArgs<T> a) {
  const SwGeo& g = a.g;
  const int64_t n = (int64_t)g.NY * g.NX;
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int NX = g.NX, NY = g.NY;
  const int j = (int)(k / NX), i = (int)(k - (int64_t)j * NX);
  const int js = a.fill ? wsrc_j(g, j) : j, is = a.fill ? wsrc_i(g, i) : i;
  const int64_t q = (int64_t)js * NX + is;
  // du.x: vortexforce (V = U.y, s = yshift, s2 = xshift, sign +1)
  T vx = vforce<T>(a.method, a.Uy, a.w, a.ovy[q], q, NX, 1, T(1), n);
  // du.y: vortexforce (V = U.x, s = xshift, s2 = yshift, sign -1)
  T vy = vforce<T>(a.method, a.Ux, a.w, a.ovx[q], q, 1, NX, T(-1), n);
  const T ff = T(a.f);
  if (js < NY - 1 && is >= 1 && is < NX - 1)
    vx = vx + ff * (((a.Uy[q - 1] + a.Uy[q + NX - 1]) + a.Uy[q]) + a.Uy[q + NX]);
  if (js >= 1 && js < NY - 1 && is < NX - 1)
    vy = vy - ff * (((a.Ux[q - NX] + a.Ux[q - NX + 1]) + a.Ux[q]) + a.Ux[q + 1]);
  if (a.ke) {
    const T* phis[2] = {a.ke, a.p};
    for (int t = 0; t < 2; t++) {
      const T* phi = phis[t];
      if (is >= 1) vx = vx - (phi[q] - phi[q - 1]) * T(a.mx[q]);
      if (js >= 1) vy = vy - (phi[q] - phi[q - NX]) * T(a.my[q]);
    }
  }
  double c[3] = {0.0, 0.0, 0.0};
  const int nc = a.integ < 0 ? 0 : rk_coefs(a.integ, a.stage, *a.dt, c);
  stage_out<T>(a.tx, nc, c, k, vx);
  stage_out<T>(a.ty, nc, c, k, vy);
}

template <class T>
struct DiagArgs {
  SwGeo g;
  int method;
  const T* ux;
  const T* uy;
  const T* h;
  const double* slip;
  const int8_t* okx;
  const int8_t* oky;
  const int8_t* msk;
  double sx, sy, goa, hb;
  T* Ux;
  T* Uy;
  T* w;
  T* ke;  // nullable (qgrsw)
  T* p;   // nullable (qgrsw)
};

template <class T>
__global__ __launch_bounds__(256) void k_sw_diag(DiagArgs<T> a) {
  const SwGeo& g = a.g;
  const int64_t n = (int64_t)g.NY * g.NX;
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int NX = g.NX;
  const int j = (int)(k / NX), i = (int)(k - (int64_t)j * NX);
  const T sx = T(a.sx), sy = T(a.sy);
  // sharp (operators.py:59-64), at every point
  a.Ux[k] = a.ux[k] * sx;
  a.Uy[k] = a.uy[k] * sy;
  // compute_pressure (operators.py:110-111), at every point, unfilled
  if (a.p) a.p[k] = T(a.goa) * (a.h[k] + T(a.hb));
  // vorticity and kinetic energy are filled: the source's value
  const int js = wsrc_j(g, j), is = wsrc_i(g, i);
  const int64_t q = (int64_t)js * NX + is;
  // compute_vorticity (operators.py:67-77): row 0 keeps its old value; it
  // has images only through a periodic x (y closed), which its own thread
  // writes
  if (js >= 1 || q == k) {
    T v = js >= 1 ? -(a.ux[q] - a.ux[q - NX]) : a.w[q];
    if (is >= 1) v = v + (a.uy[q] - a.uy[q - 1]);
    const T wv = T((double)v * a.slip[q]);
    a.w[k] = wv;
    if (js == 0 && g.xper) {
      if (i >= 3 && i < 6) a.w[k + g.nx] = wv;
      if (i >= g.nx && i < g.nx + 3) a.w[k - g.nx] = wv;
    }
  }
  if (a.ke) {
    // compute_kinetic_energy, "weno"-family methods (operators.py:93-101):
    // ke = 0; ke += U.x-term; ke += U.y-term; ke *= msk * 0.5
    T v = T(0);
    const int ox = a.okx[q], oy = a.oky[q];
    if (ox > 0) v = v + iprod<T>(a.method, a.ux, sx, ox, q, 1, n);
    if (oy > 0) v = v + iprod<T>(a.method, a.uy, sy, oy, q, NX, n);
    a.ke[k] = T((double)v * ((double)a.msk[q] * 0.5));
  }
}

// qg_projection's right-hand side (operators.py:176-180, 194-201): pv =
// vorticity(du) * slip (row 0 keeping its old pv), then pv += c2v(dh *
// scale) on [1:, 1:], pv *= mskv.  Every point from its own index.
template <class T>
__global__ __launch_bounds__(256) void k_qg_pv(SwGeo g, const T* __restrict__ dux,
                                               const T* __restrict__ duy,
                                               const T* __restrict__ dh,
                                               const double* __restrict__ slip,
                                               const int8_t* __restrict__ mv, double scale,
                                               T* __restrict__ pv) {
  const int64_t n = (int64_t)g.NY * g.NX;
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int NX = g.NX;
  const int j = (int)(k / NX), i = (int)(k - (int64_t)j * NX);
  T v = j >= 1 ? -(dux[k] - dux[k - NX]) : pv[k];
  if (i >= 1) v = v + (duy[k] - duy[k - 1]);
  T val = T((double)v * slip[k]);
  if (j >= 1 && i >= 1) {
    const T sc = T(scale);
    const T s = T(0.25) * ((((dh[k - NX - 1] * sc) + (dh[k - 1] * sc)) + (dh[k - NX] * sc)) +
                           (dh[k] * sc));
    val = val + s;
  }
  pv[k] = val * T(mv[k]);
}

template <class T>
struct ProjArgs {
  SwGeo g;
  int integ, stage;
  const T* psi;
  const int8_t* mx;
  const int8_t* my;
  const int8_t* mv;
  const int8_t* msk;
  double scale;  // f0 * area / g
  const T* pux;  // tendencies before the projection (their unsliced
  const T* puy;  // entries stay)
  const T* ph;
  Terms<T> tx, ty, th;
  const double* dt;
};

// after the Helmholtz solve (operators.py:181-183, 204-211, equations.py
// 131): dh = verticestocenters(psi * scale) (cells without a wet vertex:
// (1/0) * 0 = NaN as in numpy) times msk, du = perpgrad(psi) (covariant),
// fill(du, dh), then the RK addto of u and h.  Every point takes its
// source's values (the fill); the rows and columns the slices skip keep the
// pre-projection tendency.
template <class T>
__global__ __launch_bounds__(256) void k_qg_project(ProjArgs<T> a) {
  const SwGeo& g = a.g;
  const int64_t n = (int64_t)g.NY * g.NX;
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int NX = g.NX, NY = g.NY;
  const int j = (int)(k / NX), i = (int)(k - (int64_t)j * NX);
  const int js = wsrc_j(g, j), is = wsrc_i(g, i);
  const int64_t q = (int64_t)js * NX + is;
  const T* psi = a.psi;
  T dh;
  if (js < NY - 1 && is < NX - 1) {
    const int coef = (((int)a.mv[q] + (int)a.mv[q + NX]) + (int)a.mv[q + 1]) + (int)a.mv[q + NX + 1];
    const T sc = T(a.scale);
    const T s = (((psi[q] * sc) + (psi[q + NX] * sc)) + (psi[q + 1] * sc)) + (psi[q + NX + 1] * sc);
    dh = T((1.0 / (double)coef) * (double)s);
  } else {
    dh = a.ph[q];
  }
  dh = dh * T(a.msk[q]);
  const T dux = js < NY - 1 ? (-(psi[q + NX] - psi[q])) * T(a.mx[q]) : a.pux[q];
  const T duy = is < NX - 1 ? (psi[q + 1] - psi[q]) * T(a.my[q]) : a.puy[q];
  double c[3] = {0.0, 0.0, 0.0};
  const int nc = a.integ < 0 ? 0 : rk_coefs(a.integ, a.stage, *a.dt, c);
  stage_out<T>(a.tx, nc, c, k, dux);
  stage_out<T>(a.ty, nc, c, k, duy);
  stage_out<T>(a.th, nc, c, k, dh);
}

}  // namespace f2d

using namespace f2d;

#define F2D_SW_DISPATCH(dtype, ...)         \
  do {                                      \
    if ((dtype) == F2D_F64) {               \
      using T = double;                     \
      __VA_ARGS__;                          \
    } else if ((dtype) == F2D_F32) {        \
      using T = float;                      \
      __VA_ARGS__;                          \
    } else {                                \
      throw ::f2d::Error{"unknown dtype"};  \
    }                                       \
  } while (0)

static SwGeo sw_geo(const f2d_grid* g) {
  F2D_REQUIRE(g != nullptr, "grid descriptor is NULL");
  F2D_REQUIRE(g->halowidth == 3, "halowidth must be 3");
  F2D_REQUIRE(g->ny > 0 && g->nx > 0 && g->ny < (1 << 24) && g->nx < (1 << 24),
              "grid is empty or too large");
  SwGeo s;
  s.ny = (int)g->ny;
  s.nx = (int)g->nx;
  s.NY = s.ny + 6;
  s.NX = s.nx + 6;
  s.xper = g->xperiodic;
  s.yper = g->yperiodic;
  return s;
}

template <class T>
static Terms<T> terms(const f2d_sw_stage* st, int f) {
  Terms<T> t{};
  if (!st) return t;
  t.p0 = (const T*)st->prev[0][f];
  t.p1 = (const T*)st->prev[1][f];
  t.out = (T*)st->out[f];
  t.y = (T*)st->y[f];
  return t;
}

static int stage_check(const f2d_sw_stage* st, int nf) {
  if (!st || st->integrator < 0) return -1;
  F2D_REQUIRE(st->integrator <= F2D_ENRK3, "unknown integrator");
  const int nst = st->integrator == F2D_EF ? 1 : 3;
  F2D_REQUIRE(st->stage >= 0 && st->stage < nst, "bad stage");
  F2D_REQUIRE(st->dt, "dt is NULL");
  for (int f = 0; f < nf; f++) {
    if (st->stage >= 1) F2D_REQUIRE(st->prev[0][f], "earlier tendency missing");
    if (st->stage >= 2) F2D_REQUIRE(st->prev[1][f], "earlier tendency missing");
  }
  return st->integrator;
}

static unsigned nblk(const SwGeo& g) {
  return (unsigned)(((int64_t)g.NY * g.NX + 255) / 256);
}

extern "C" {

int f2d_sw_momentum(int dtype, const f2d_grid* g, int method, const void* Ux, const void* Uy,
                    const void* omega, const int8_t* ovx, const int8_t* ovy, double f,
                    const void* ke, const void* p, const int8_t* mskx, const int8_t* msky,
                    int fill, const f2d_sw_stage* st, void* stream) {
  return guarded([&] {
    const SwGeo geo = sw_geo(g);
    F2D_REQUIRE(method >= F2D_WENO && method <= F2D_CWENO, "unknown method");
    F2D_REQUIRE(Ux && Uy && omega && ovx && ovy, "NULL argument");
    F2D_REQUIRE((ke == nullptr) == (p == nullptr), "ke and p go together");
    F2D_REQUIRE(!ke || (mskx && msky), "masks are required with the gradient terms");
    const int integ = stage_check(st, 2);
    F2D_SW_DISPATCH(dtype, {
      MomArgs<T> a{};
      a.g = geo;
      a.method = method;
      a.fill = fill;
      a.integ = integ;
      a.stage = st ? st->stage : 0;
      a.Ux = (const T*)Ux;
      a.Uy = (const T*)Uy;
      a.w = (const T*)omega;
      a.ovx = ovx;
      a.ovy = ovy;
      a.f = f;
      a.ke = (const T*)ke;
      a.p = (const T*)p;
      a.mx = mskx;
      a.my = msky;
      a.tx = terms<T>(st, 0);
      a.ty = terms<T>(st, 1);
      a.dt = st ? st->dt : nullptr;
      k_sw_momentum<T><<<nblk(geo), 256, 0, S(stream)>>>(a);
    });
  });
}

int f2d_sw_diag(int dtype, const f2d_grid* g, int method, const void* ux, const void* uy,
                const void* h, const double* slipcoef, const int8_t* okx, const int8_t* oky,
                const int8_t* msk, double g_over_area, double hb, void* Ux, void* Uy,
                void* omega, void* ke, void* p, void* stream) {
  return guarded([&] {
    const SwGeo geo = sw_geo(g);
    F2D_REQUIRE(ux && uy && slipcoef && Ux && Uy && omega, "NULL argument");
    F2D_REQUIRE(!ke || (method >= F2D_WENO && method <= F2D_CWENO && okx && oky && msk),
                "kinetic energy needs a flux method (not classic), orders and msk");
    F2D_REQUIRE(!p || h, "pressure needs h");
    F2D_SW_DISPATCH(dtype, {
      DiagArgs<T> a{};
      a.g = geo;
      a.method = method;
      a.ux = (const T*)ux;
      a.uy = (const T*)uy;
      a.h = (const T*)h;
      a.slip = slipcoef;
      a.okx = okx;
      a.oky = oky;
      a.msk = msk;
      a.sx = 1.0 / (g->dx * g->dx);  // sharp's 1/dx**2, 1/dy**2
      a.sy = 1.0 / (g->dy * g->dy);
      a.goa = g_over_area;
      a.hb = hb;
      a.Ux = (T*)Ux;
      a.Uy = (T*)Uy;
      a.w = (T*)omega;
      a.ke = (T*)ke;
      a.p = (T*)p;
      k_sw_diag<T><<<nblk(geo), 256, 0, S(stream)>>>(a);
    });
  });
}

int f2d_qg_pv(int dtype, const f2d_grid* g, const void* dux, const void* duy, const void* dh,
              const double* slipcoef, const int8_t* mskv, double scale, void* pv,
              void* stream) {
  return guarded([&] {
    const SwGeo geo = sw_geo(g);
    F2D_REQUIRE(dux && duy && dh && slipcoef && mskv && pv, "NULL argument");
    F2D_SW_DISPATCH(dtype, k_qg_pv<T><<<nblk(geo), 256, 0, S(stream)>>>(
                               geo, (const T*)dux, (const T*)duy, (const T*)dh, slipcoef,
                               mskv, scale, (T*)pv));
  });
}

int f2d_qg_project(int dtype, const f2d_grid* g, const void* psi, const int8_t* mskx,
                   const int8_t* msky, const int8_t* mskv, const int8_t* msk, double scale,
                   const void* pre_ux, const void* pre_uy, const void* pre_h,
                   const f2d_sw_stage* st, void* stream) {
  return guarded([&] {
    const SwGeo geo = sw_geo(g);
    F2D_REQUIRE(psi && mskx && msky && mskv && msk && pre_ux && pre_uy && pre_h,
                "NULL argument");
    F2D_REQUIRE(st, "stage terms are NULL");
    for (int f = 0; f < 3; f++) {
      const void* pre = f == 0 ? pre_ux : (f == 1 ? pre_uy : pre_h);
      F2D_REQUIRE(st->out[f] != pre && st->y[f] != pre,
                  "the projected tendency must not overwrite the unprojected one");
    }
    const int integ = stage_check(st, 3);
    F2D_SW_DISPATCH(dtype, {
      ProjArgs<T> a{};
      a.g = geo;
      a.integ = integ;
      a.stage = st->stage;
      a.psi = (const T*)psi;
      a.mx = mskx;
      a.my = msky;
      a.mv = mskv;
      a.msk = msk;
      a.scale = scale;
      a.pux = (const T*)pre_ux;
      a.puy = (const T*)pre_uy;
      a.ph = (const T*)pre_h;
      a.tx = terms<T>(st, 0);
      a.ty = terms<T>(st, 1);
      a.th = terms<T>(st, 2);
      a.dt = st->dt;
      k_qg_project<T><<<nblk(geo), 256, 0, S(stream)>>>(a);
    });
  });
}

}  // extern "C"
