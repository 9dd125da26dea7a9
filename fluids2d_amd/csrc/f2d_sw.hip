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
// high side), as the flat loops of f2d_stencil.hip.  Flat indices are 32-bit
// here: the entry points refuse grids of 2^31 - 4 (nx + 6) points or more.
template <bool WRAP>
__device__ __forceinline__ int ix(int i, int n) {
  return WRAP ? (i < 0 ? i + n : (i >= n ? i - n : i)) : i;
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

// Workgroups of 64 x 4 points (columns x rows) over the padded array, in
// tiles numbered row-major.  The hardware places workgroup b on XCD b % 8;
// workgroup b takes the (b / 8)-th tile of a contiguous band per XCD, so the
// rows a tile's stencils share with the tiles above and below it are mostly
// fetched into the same XCD's L2.
constexpr int TW = 64, TH = 4;

__device__ __forceinline__ bool tile_point(const SwGeo& g, int& i, int& j) {
  const int gx = (g.NX + TW - 1) / TW;
  const int nb = gridDim.x, b = blockIdx.x;
  const int x = b & 7, q8 = nb >> 3, rem = nb & 7;
  const int t = x * q8 + (x < rem ? x : rem) + (b >> 3);
  const int ty = t / gx, tx = t - ty * gx;
  i = tx * TW + threadIdx.x;
  j = ty * TH + threadIdx.y;
  return i < g.NX && j < g.NY;
}

// Whether every flat index a stencil of source row js reaches (rows js - 2
// to js + 3) lies inside the array: then no wrap is needed.
__device__ __forceinline__ bool no_wrap(const SwGeo& g, int js) {
  return js >= 2 && js <= g.NY - 4;
}

// vortexforce_on_interval at flat index i (weno.py:370-384): V the
// transverse contravariant velocity, q the vertex vorticity, o the order
// V read from the array, or (SC) formed as the array times vs: the
// contravariant velocity sharp(u) = u * (1/dx**2, 1/dy**2) as k_sharp
// stores it
template <class T, bool SC>
__device__ __forceinline__ T vel(const T* V, T vs, int x) {
  return SC ? V[x] * vs : V[x];
}

template <class T, int M, bool W, bool SC>
__device__ __forceinline__ T vforce(const T* V, T vs, const T* q, int o, int i, int s, int s2,
                                    T sign, int n) {
  if (o <= 0) return T(0);
  const T Vm = T(0.25) * (((vel<T, SC>(V, vs, i) + vel<T, SC>(V, vs, ix<W>(i + s, n))) +
                           vel<T, SC>(V, vs, ix<W>(i - s2, n))) +
                          vel<T, SC>(V, vs, ix<W>(i + s - s2, n)));
  T f;
  if (o > 4)
    f = f5<T>(M, Vm, q[ix<W>(i - 2 * s, n)], q[ix<W>(i - s, n)], q[i], q[ix<W>(i + s, n)],
              q[ix<W>(i + 2 * s, n)], q[ix<W>(i + 3 * s, n)]);
  else if (o > 2)
    f = f3<T>(M, Vm, q[ix<W>(i - s, n)], q[i], q[ix<W>(i + s, n)], q[ix<W>(i + 2 * s, n)]);
  else
    f = f1<T>(M, Vm, q[i], q[ix<W>(i + s, n)]);
  return (sign * f) * Vm;
}

// innerproduct_on_interval term at flat index i (weno.py:391-403), with the
// contravariant velocity U = u * sc formed from u as sharp stores it
template <class T, int M, bool W>
__device__ __forceinline__ T iprod(const T* u, T sc, int o, int i, int s, int n) {
  const T Um = T(0.5) * (u[i] * sc + u[ix<W>(i + s, n)] * sc);
  T f;
  if (o > 4)
    f = f5<T>(M, Um, u[ix<W>(i - 2 * s, n)], u[ix<W>(i - s, n)], u[i], u[ix<W>(i + s, n)],
              u[ix<W>(i + 2 * s, n)], u[ix<W>(i + 3 * s, n)]);
  else if (o > 2)
    f = f3<T>(M, Um, u[ix<W>(i - s, n)], u[i], u[ix<W>(i + s, n)], u[ix<W>(i + 2 * s, n)]);
  else
    f = f1<T>(M, Um, u[i], u[ix<W>(i + s, n)]);
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
  const T* yin;  // nullable: y = yin + (...) instead of y += (...)
};

template <class T>
__device__ __forceinline__ void stage_out(const Terms<T>& t, int nc, const double c[3], int k,
                                          T ds) {
  if (t.out) t.out[k] = ds;
  if (t.y && nc > 0) {
    const T x0 = (nc == 1) ? ds : t.p0[k];
    const T x1 = (nc == 2) ? ds : (nc > 2 ? t.p1[k] : T(0));
    const T x2 = (nc == 3) ? ds : T(0);
    t.y[k] = (t.yin ? t.yin[k] : t.y[k]) + rk_combine<T>(nc, c, x0, x1, x2);
  }
}

template <class T>
struct MomArgs {
  SwGeo g;
  int fill, integ, stage;
  const T* Ux;  // U, or with the FU kernels u (U = u * (sx, sy))
  const T* Uy;
  double sx, sy;
  const T* w;
  const int8_t* ovx;
  const int8_t* ovy;
  double f;  // addcoriolis coefficient f0 * area * 0.25
  const T* ke;  // nullable (qgrsw: no gradient terms)
  const T* p;   // p, or with the FU kernels h (p = goa * (h + hb))
  double goa, hb;
  const int8_t* mx;
  const int8_t* my;
  Terms<T> tx, ty;
  const double* dt;
  double* dt_log;  // receives *dt (one thread), nullable
};

// vortex force of both components at source point q (row js, column is)
template <class T, int M, bool W, bool FU>
__device__ __forceinline__ void vforce2(const MomArgs<T>& a, int q, int n, T& vx, T& vy) {
  const int NX = a.g.NX;
  // du.x: vortexforce (V = U.y, s = yshift, s2 = xshift, sign +1)
  vx = vforce<T, M, W, FU>(a.Uy, T(a.sy), a.w, a.ovy[q], q, NX, 1, T(1), n);
  // du.y: vortexforce (V = U.x, s = xshift, s2 = yshift, sign -1)
  vy = vforce<T, M, W, FU>(a.Ux, T(a.sx), a.w, a.ovx[q], q, 1, NX, T(-1), n);
}

// pressure at x: read, or (FU) compute_pressure's (g/area) * (h + hb)
template <class T, bool FU>
__device__ __forceinline__ T pres(const MomArgs<T>& a, int x) {
  return FU ? T(a.goa) * (a.p[x] + T(a.hb)) : a.p[x];
}

// FU: U = sharp(u) and p = (g/area)(h + hb) formed from u and h, with the
// arithmetic of k_sharp and k_pressure, instead of read
// B: the batched instantiation (a = tab[blockIdx.y]: one member of a
// recorded member batch, launch_rec)
template <class T, int M, bool FU, bool B = false>
__global__ __launch_bounds__(256) void k_sw_momentum(MomArgs<T> a, const MomArgs<T>* tab) {
  if (B) a = tab[blockIdx.y];
  const SwGeo& g = a.g;
  int i, j;
  if (!tile_point(g, i, j)) return;
  const int NX = g.NX, NY = g.NY, n = NY * NX;
  const int k = j * NX + i;
  const int js = a.fill ? wsrc_j(g, j) : j, is = a.fill ? wsrc_i(g, i) : i;
  const int q = js * NX + is;
  if (a.dt_log && k == 0) *a.dt_log = *a.dt;
  T vx, vy;
  if (no_wrap(g, js))
    vforce2<T, M, false, FU>(a, q, n, vx, vy);
  else
    vforce2<T, M, true, FU>(a, q, n, vx, vy);
  const T ff = T(a.f);
  const T sx = T(a.sx), sy = T(a.sy);
  if (js < NY - 1 && is >= 1 && is < NX - 1)
    vx = vx + ff * (((vel<T, FU>(a.Uy, sy, q - 1) + vel<T, FU>(a.Uy, sy, q + NX - 1)) +
                     vel<T, FU>(a.Uy, sy, q)) + vel<T, FU>(a.Uy, sy, q + NX));
  if (js >= 1 && js < NY - 1 && is < NX - 1)
    vy = vy - ff * (((vel<T, FU>(a.Ux, sx, q - NX) + vel<T, FU>(a.Ux, sx, q - NX + 1)) +
                     vel<T, FU>(a.Ux, sx, q)) + vel<T, FU>(a.Ux, sx, q + 1));
  if (a.ke) {
    // addgrad(ke), then addgrad(p)
    if (is >= 1) vx = vx - (a.ke[q] - a.ke[q - 1]) * T(a.mx[q]);
    if (js >= 1) vy = vy - (a.ke[q] - a.ke[q - NX]) * T(a.my[q]);
    if (is >= 1) vx = vx - (pres<T, FU>(a, q) - pres<T, FU>(a, q - 1)) * T(a.mx[q]);
    if (js >= 1) vy = vy - (pres<T, FU>(a, q) - pres<T, FU>(a, q - NX)) * T(a.my[q]);
  }
  double c[3] = {0.0, 0.0, 0.0};
  const int nc = a.integ < 0 ? 0 : rk_coefs(a.integ, a.stage, *a.dt, c);
  stage_out<T>(a.tx, nc, c, k, vx);
  stage_out<T>(a.ty, nc, c, k, vy);
}

template <class T>
struct DiagArgs {
  SwGeo g;
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

// compute_kinetic_energy, "weno"-family methods (operators.py:93-101):
// ke = 0; ke += U.x-term; ke += U.y-term (then ke *= msk * 0.5)
template <class T, int M, bool W>
__device__ __forceinline__ T ke_sum(const DiagArgs<T>& a, T sx, T sy, int q, int n) {
  T v = T(0);
  const int ox = a.okx[q], oy = a.oky[q];
  if (ox > 0) v = v + iprod<T, M, W>(a.ux, sx, ox, q, 1, n);
  if (oy > 0) v = v + iprod<T, M, W>(a.uy, sy, oy, q, a.g.NX, n);
  return v;
}

template <class T, int M, bool B = false>
__global__ __launch_bounds__(256) void k_sw_diag(DiagArgs<T> a, const DiagArgs<T>* tab) {
  if (B) a = tab[blockIdx.y];
  const SwGeo& g = a.g;
  int i, j;
  if (!tile_point(g, i, j)) return;
  const int NX = g.NX, n = g.NY * NX;
  const int k = j * NX + i;
  const T sx = T(a.sx), sy = T(a.sy);
  // sharp (operators.py:59-64), at every point
  if (a.Ux) {
    a.Ux[k] = a.ux[k] * sx;
    a.Uy[k] = a.uy[k] * sy;
  }
  // compute_pressure (operators.py:110-111), at every point, unfilled
  if (a.p) a.p[k] = T(a.goa) * (a.h[k] + T(a.hb));
  // vorticity and kinetic energy are filled: the source's value
  const int js = wsrc_j(g, j), is = wsrc_i(g, i);
  const int q = js * NX + is;
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
    const T v = no_wrap(g, js) ? ke_sum<T, M, false>(a, sx, sy, q, n)
                               : ke_sum<T, M, true>(a, sx, sy, q, n);
    a.ke[k] = T((double)v * ((double)a.msk[q] * 0.5));
  }
}

// qg_projection's right-hand side (operators.py:176-180, 194-201): pv =
// vorticity(du) * slip (row 0 keeping its old pv), then pv += c2v(dh *
// scale) on [1:, 1:], pv *= mskv.  Every point from its own index.
template <class T>
struct PvArgs {
  SwGeo g;
  const T* dux;
  const T* duy;
  const T* dh;
  const double* slip;
  const int8_t* mv;
  double scale;
  T* pv;
};

template <class T, bool B = false>
__global__ __launch_bounds__(256) void k_qg_pv(PvArgs<T> a, const PvArgs<T>* tab) {
  if (B) a = tab[blockIdx.y];
  const SwGeo& g = a.g;
  const T* __restrict__ dux = a.dux;
  const T* __restrict__ duy = a.duy;
  const T* __restrict__ dh = a.dh;
  const double* __restrict__ slip = a.slip;
  const int8_t* __restrict__ mv = a.mv;
  const double scale = a.scale;
  T* __restrict__ pv = a.pv;
  int i, j;
  if (!tile_point(g, i, j)) return;
  const int NX = g.NX;
  const int k = j * NX + i;
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
template <class T, bool B = false>
__global__ __launch_bounds__(256) void k_qg_project(ProjArgs<T> a, const ProjArgs<T>* tab) {
  if (B) a = tab[blockIdx.y];
  const SwGeo& g = a.g;
  int i, j;
  if (!tile_point(g, i, j)) return;
  const int NX = g.NX, NY = g.NY;
  const int k = j * NX + i;
  const int js = wsrc_j(g, j), is = wsrc_i(g, i);
  const int q = js * NX + is;
  const T* psi = a.psi;
  T dh;
  if (js < NY - 1 && is < NX - 1) {
    const int coef = (((int)a.mv[q] + (int)a.mv[q + NX]) + (int)a.mv[q + 1]) + (int)a.mv[q + NX + 1];
    const T sc = T(a.scale);
    const T s = (((psi[q] * sc) + (psi[q + NX] * sc)) + (psi[q + 1] * sc)) + (psi[q + NX + 1] * sc);
    dh = T(inv_count(coef) * (double)s);
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
  F2D_REQUIRE((g->ny + 6) * (g->nx + 6) + 4 * (g->nx + 6) < ((int64_t)1 << 31),
              "grid too large for the fused shallow-water kernels (2^31 points)");
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
  t.yin = (const T*)st->y_in[f];
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

// launch shape of the tiled kernels: one workgroup of TW x TH threads per
// tile (tile_point)
static unsigned nblk(const SwGeo& g) {
  return (unsigned)(((g.NX + TW - 1) / TW) * ((g.NY + TH - 1) / TH));
}
static const dim3 TILE_BLOCK(TW, TH);

template <class T, int M, bool FU>
static void launch_momentum(const MomArgs<T>& a, const SwGeo& geo, void* stream) {
  launch_rec(k_sw_momentum<T, M, FU, false>, k_sw_momentum<T, M, FU, true>, nblk(geo),
             TILE_BLOCK, 0, S(stream), a);
}

template <class T, bool FU>
static void launch_momentum(const MomArgs<T>& a, const SwGeo& geo, int method, void* stream) {
  switch (method) {
    case F2D_WENO: launch_momentum<T, F2D_WENO, FU>(a, geo, stream); break;
    case F2D_UPWIND: launch_momentum<T, F2D_UPWIND, FU>(a, geo, stream); break;
    case F2D_CENTERED: launch_momentum<T, F2D_CENTERED, FU>(a, geo, stream); break;
    default: launch_momentum<T, F2D_CWENO, FU>(a, geo, stream); break;
  }
}

template <class T, int M>
static void launch_diag(const DiagArgs<T>& a, const SwGeo& geo, void* stream) {
  launch_rec(k_sw_diag<T, M, false>, k_sw_diag<T, M, true>, nblk(geo), TILE_BLOCK, 0,
             S(stream), a);
}

template <class T>
static void launch_diag(const DiagArgs<T>& a, int method, const SwGeo& geo, void* stream) {
  switch (method) {
    case F2D_WENO: launch_diag<T, F2D_WENO>(a, geo, stream); break;
    case F2D_UPWIND: launch_diag<T, F2D_UPWIND>(a, geo, stream); break;
    case F2D_CENTERED: launch_diag<T, F2D_CENTERED>(a, geo, stream); break;
    default: launch_diag<T, F2D_CWENO>(a, geo, stream); break;
  }
}

// the momentum entry points: from_u selects U = sharp(u), p from h
static int momentum_impl(int dtype, const f2d_grid* g, int method, const void* Ux,
                         const void* Uy, const void* omega, const int8_t* ovx,
                         const int8_t* ovy, double f, const void* ke, const void* p,
                         double goa, double hb, const int8_t* mskx, const int8_t* msky,
                         int fill, const f2d_sw_stage* st, void* stream, bool from_u) {
  return guarded_rec([&] {
    const SwGeo geo = sw_geo(g);
    F2D_REQUIRE(method >= F2D_WENO && method <= F2D_CWENO, "unknown method");
    F2D_REQUIRE(Ux && Uy && omega && ovx && ovy, "NULL argument");
    F2D_REQUIRE((ke == nullptr) == (p == nullptr), "ke and p (or h) go together");
    F2D_REQUIRE(!ke || (mskx && msky), "masks are required with the gradient terms");
    const int integ = stage_check(st, 2);
    F2D_REQUIRE(!st || !st->dt_log || st->dt, "dt_log needs dt");
    if (from_u && st)  // u is read around each point: it cannot be updated in place
      for (int f = 0; f < 2; f++)
        F2D_REQUIRE(st->y[f] != Ux && st->y[f] != Uy,
                    "the momentum from u must update into another array (y_in = u)");
    F2D_SW_DISPATCH(dtype, {
      MomArgs<T> a{};
      a.g = geo;
      a.fill = fill;
      a.integ = integ;
      a.stage = st ? st->stage : 0;
      a.Ux = (const T*)Ux;
      a.Uy = (const T*)Uy;
      a.sx = 1.0 / (g->dx * g->dx);  // sharp's 1/dx**2, 1/dy**2
      a.sy = 1.0 / (g->dy * g->dy);
      a.w = (const T*)omega;
      a.ovx = ovx;
      a.ovy = ovy;
      a.f = f;
      a.ke = (const T*)ke;
      a.p = (const T*)p;
      a.goa = goa;
      a.hb = hb;
      a.mx = mskx;
      a.my = msky;
      a.tx = terms<T>(st, 0);
      a.ty = terms<T>(st, 1);
      a.dt = st ? st->dt : nullptr;
      a.dt_log = st ? st->dt_log : nullptr;
      if (from_u)
        launch_momentum<T, true>(a, geo, method, stream);
      else
        launch_momentum<T, false>(a, geo, method, stream);
    });
  });
}

extern "C" {

int f2d_sw_momentum(int dtype, const f2d_grid* g, int method, const void* Ux, const void* Uy,
                    const void* omega, const int8_t* ovx, const int8_t* ovy, double f,
                    const void* ke, const void* p, const int8_t* mskx, const int8_t* msky,
                    int fill, const f2d_sw_stage* st, void* stream) {
  return momentum_impl(dtype, g, method, Ux, Uy, omega, ovx, ovy, f, ke, p, 0.0, 0.0, mskx,
                       msky, fill, st, stream, false);
}

int f2d_sw_momentum_u(int dtype, const f2d_grid* g, int method, const void* ux,
                      const void* uy, const void* omega, const int8_t* ovx, const int8_t* ovy,
                      double f, const void* ke, const void* h, double g_over_area, double hb,
                      const int8_t* mskx, const int8_t* msky, int fill,
                      const f2d_sw_stage* st, void* stream) {
  return momentum_impl(dtype, g, method, ux, uy, omega, ovx, ovy, f, ke, h, g_over_area, hb,
                       mskx, msky, fill, st, stream, true);
}

int f2d_sw_diag(int dtype, const f2d_grid* g, int method, const void* ux, const void* uy,
                const void* h, const double* slipcoef, const int8_t* okx, const int8_t* oky,
                const int8_t* msk, double g_over_area, double hb, void* Ux, void* Uy,
                void* omega, void* ke, void* p, void* stream) {
  return guarded_rec([&] {
    const SwGeo geo = sw_geo(g);
    F2D_REQUIRE(ux && uy && slipcoef && omega, "NULL argument");
    F2D_REQUIRE((Ux == nullptr) == (Uy == nullptr), "Ux and Uy go together");
    F2D_REQUIRE(!ke || (method >= F2D_WENO && method <= F2D_CWENO && okx && oky && msk),
                "kinetic energy needs a flux method (not classic), orders and msk");
    F2D_REQUIRE(!p || h, "pressure needs h");
    F2D_SW_DISPATCH(dtype, {
      DiagArgs<T> a{};
      a.g = geo;
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
      launch_diag<T>(a, ke ? method : F2D_WENO, geo, stream);
    });
  });
}

int f2d_qg_pv(int dtype, const f2d_grid* g, const void* dux, const void* duy, const void* dh,
              const double* slipcoef, const int8_t* mskv, double scale, void* pv,
              void* stream) {
  return guarded_rec([&] {
    const SwGeo geo = sw_geo(g);
    F2D_REQUIRE(dux && duy && dh && slipcoef && mskv && pv, "NULL argument");
    F2D_SW_DISPATCH(dtype, {
      const PvArgs<T> a{geo, (const T*)dux, (const T*)duy, (const T*)dh, slipcoef, mskv,
                        scale, (T*)pv};
      launch_rec(k_qg_pv<T, false>, k_qg_pv<T, true>, nblk(geo), TILE_BLOCK, 0, S(stream), a);
    });
  });
}

int f2d_qg_project(int dtype, const f2d_grid* g, const void* psi, const int8_t* mskx,
                   const int8_t* msky, const int8_t* mskv, const int8_t* msk, double scale,
                   const void* pre_ux, const void* pre_uy, const void* pre_h,
                   const f2d_sw_stage* st, void* stream) {
  return guarded_rec([&] {
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
      launch_rec(k_qg_project<T, false>, k_qg_project<T, true>, nblk(geo), TILE_BLOCK, 0,
                 S(stream), a);
    });
  });
}

}  // extern "C"
