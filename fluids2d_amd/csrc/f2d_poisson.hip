// f2d_poisson.hip -- masked Poisson/Helmholtz solve on the vertex grid:
// geometric-multigrid-preconditioned conjugate gradients.
//
// Replaces Poisson2D (elliptic.py:71-87): the reference assembles the masked
// 5-point operator (elliptic.py:114-195) and factorises it with SuperLU.
// Here the operator is applied matrix-free.  With K = -A (SPD):
//   (K x)_P = m_P [ d x_P - cx (x_E + x_W) - cy (x_N + x_S) ],
//   cx = dy/dx, cy = dx/dy, d = 2(cx + cy) + maindiag,
// vectors vanishing at dry points (Dirichlet), neighbours wrapping in
// periodic directions.  We solve K x = -b.
//
// Preconditioner: one symmetric V-cycle.  Coarse vertex (J, I) sits on fine
// vertex (2J, 2I); the coarse mask is injected, the coarse operator is the
// same 5-point form with cx/4, cy/4 (the maindiag term is unchanged),
// prolongation is bilinear and restriction is (1/4) of its transpose (full
// weighting).  Smoother: red-black Gauss-Seidel, pre-smoothing in the order
// (red, black)^nu and post-smoothing (black, red)^nu, so the V-cycle is a
// symmetric positive-definite operator.
//
// Kernels of CG iteration k (level 0 = the fine grid, 32x32 tiles):
//   k_down<CG>   alpha from the previous k_pupdate's partial sums;
//                r -= alpha q, x += alpha p, partial ||r||^2, then
//                pre-smooth / residual / restrict in the same LDS tile pass
//   k_down       the same smoothing/restriction on levels 1 .. Lt-1
//   k_tail       one workgroup runs the levels Lt .. L-1 entirely in LDS
//   k_up         prolongate + post-smooth on levels Lt-1 .. 1
//   k_up<FINE>   convergence test on ||r|| (partials of k_down<CG>), then
//                level-0 post-smoothing, partial (r, z), sum z, sum r
//   k_pupdate    beta from k_up<FINE>'s partials; p = z + beta p, q = K p,
//                partial (p, q), sum q
// Grid reductions never use atomics: every block writes its partial sums and
// the next kernel of the sequence (or a one-block finaliser) adds them in a
// fixed order, so results are reproducible run to run.  A converged solve
// turns the rest of the enqueued kernels into early returns.
#include <cmath>
#include <cstring>
#include <vector>

#include "f2d_capacitance.hpp"
#include "f2d_common.hpp"
#include "f2d_fft.hpp"
#include "f2d_neumann.hpp"

namespace f2d {

struct LevelDev {
  int ny, nx, stride, xper, yper;
  int64_t origin;
  double cx, cy, d, inv_d;
  const int8_t* m;
};

struct Ctl {
  double bb;      // ||g||^2 (projected for the singular operator)
  double rr;      // ||r||^2 at the last convergence test
  double relres;
  double rtol;
  double nwet;    // number of unknowns
  double mean_r;  // mean of the initial residual (singular operator)
  double rz[2];   // (r, z - mean z) of iteration k at index k & 1
  double alpha, beta, mean_z, mean_q;  // last values (diagnostics)
  int done;
  int iters;
  int x_valid;    // inner solve of a refinement: the iterate has been written
  int pad;
  // outer (refinement) control
  double obb, orr, orelres, ortol, inner_rtol;
  int odone, orefine, xcur, inner_total;
  // iterations, relres, converged, refinements of the last solve; then, since
  // the last reset: non-converged solves, max iterations, solves
  double stats[8];
};

__device__ __forceinline__ bool lw(const LevelDev& L, int& y, int& x) {
  if (x < 0 || x >= L.nx) {
    if (!L.xper) return false;
    x %= L.nx;
    if (x < 0) x += L.nx;
  }
  if (y < 0 || y >= L.ny) {
    if (!L.yper) return false;
    y %= L.ny;
    if (y < 0) y += L.ny;
  }
  return true;
}
__device__ __forceinline__ int64_t li(const LevelDev& L, int y, int x) {
  return L.origin + (int64_t)y * L.stride + x;
}

__device__ __forceinline__ bool stop(const Ctl* c, const Ctl* outer) {
  return c->done || (outer && outer->odone);
}

template <class T>
__device__ __forceinline__ T finite_or_zero(T v) {
  return isfinite(v) ? v : T(0);
}

constexpr int BS = 256;
constexpr int FC0 = 32;  // level-0 tile (points per side)
// capacitance solve: the largest boundary requested explicitly (the FP64
// gather of its values must fit a workgroup's LDS) and the largest the
// automatic choice takes (its inverse: 8192^2 FP64 = 512 MiB per handle)
constexpr int64_t CAP_MAX = 16384;
constexpr int64_t CAP_AUTO_MAX = 8192;

// Partial-sum regions (each REGION_SLOTS x nblocks doubles).
enum Region { R_INIT = 0, R_RR = 1, R_RZ = 2, R_PQ = 3, R_REF = 4, R_MEAN = 5,
              R_COUNT = 6 };
constexpr int REGION_SLOTS = 4;

// Sum K arrays of n partials (arr k at p + k*n) in a fixed order: thread t
// adds entries t, t+BS, ...; then a fixed-shape tree.  Every block that calls
// this gets bit-identical results.
template <int K>
__device__ __forceinline__ void sum_partials(const double* p, int n,
                                             double* sh, double out[K]) {
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (int i = threadIdx.x; i < n; i += BS) {
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] += p[k * n + i];
  }
#pragma unroll
  for (int k = 0; k < K; k++) out[k] = block_sum<BS>(acc[k], sh);
}

// 2-D tile of a level handled by one block: points [y0, y0+TS) x [x0, x0+TS)
template <int TS>
__device__ __forceinline__ void tile_origin(int& y0, int& x0) {
  y0 = blockIdx.y * TS;
  x0 = blockIdx.x * TS;
}

// ---------------------------------------------------------------------------
// initial residual r0 = m (g - K x0), g = -b, -c2v(omega) or given
// ---------------------------------------------------------------------------
enum GMode { G_FROM_B = 0, G_FROM_C2V = 1 };

template <class T, class TB>
struct InitArgs {
  LevelDev L;
  const TB* b;  // b (G_FROM_B) or omega cells (G_FROM_C2V), padded layout
  T* x;         // initial guess (sanitised in place), may be null (x0 = 0)
  T* r;
  Ctl* ctl;
  double* part;
  double rtol;
};

template <class TB, int GM>
__device__ __forceinline__ double g_at(const TB* b, int64_t k, int64_t P) {
  if (GM == G_FROM_B) return -(double)b[k];
  // centerstovertices, operators.py:131-132
  TB s = TB(0.25) * (((b[k - P - 1] + b[k - 1]) + b[k - P]) + b[k]);
  return -(double)s;
}

template <class T, class TB, int GM>
__global__ __launch_bounds__(BS) void k_init(InitArgs<T, TB> a) {
  __shared__ double sh[BS];
  const LevelDev& L = a.L;
  int y0, x0;
  tile_origin<FC0>(y0, x0);
  const int nb = gridDim.x * gridDim.y, bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (bid == 0 && threadIdx.x == 0) {
    Ctl* c = a.ctl;
    c->rtol = a.rtol;
    c->done = 0;
    c->iters = 0;
    c->x_valid = 0;
    c->rz[0] = c->rz[1] = 0.0;
    c->alpha = c->beta = c->mean_z = c->mean_q = 0.0;
    c->relres = 1.0;
  }
  double gg = 0.0, rs = 0.0, gs = 0.0;
  for (int e = threadIdx.x; e < FC0 * FC0; e += BS) {
    int y = y0 + e / FC0, x = x0 + e % FC0;
    if (y >= L.ny || x >= L.nx) continue;
    int64_t k = li(L, y, x);
    T rv = T(0);
    if (L.m[k]) {
      double g = g_at<TB, GM>(a.b, k, L.stride);
      T gv = T(g);
      rv = gv;
      if (a.x) {
        T xc = finite_or_zero(a.x[k]);
        a.x[k] = xc;
        T nb4[4];
        const int dy[4] = {0, 0, -1, 1}, dx[4] = {-1, 1, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; q++) {
          int yy = y + dy[q], xx = x + dx[q];
          nb4[q] = T(0);
          if (lw(L, yy, xx)) {
            int64_t kk = li(L, yy, xx);
            if (L.m[kk]) nb4[q] = finite_or_zero(a.x[kk]);
          }
        }
        T Kx = T(L.d) * xc - T(L.cx) * (nb4[0] + nb4[1]) -
               T(L.cy) * (nb4[2] + nb4[3]);
        rv = gv - Kx;
      }
      gg += (double)gv * (double)gv;
      gs += (double)gv;
      rs += (double)rv;
    }
    a.r[k] = rv;
  }
  double v0 = block_sum<BS>(gg, sh);
  double v1 = block_sum<BS>(rs, sh);
  double v2 = block_sum<BS>(gs, sh);
  if (threadIdx.x == 0) {
    a.part[bid] = v0;
    a.part[nb + bid] = v1;
    a.part[2 * nb + bid] = v2;
  }
}

// one block: ||g||^2 (projected if singular) and the mean of r0
__global__ __launch_bounds__(BS) void k_init_fin(Ctl* c, const double* part,
                                                 int nb, int singular) {
  __shared__ double sh[BS];
  double v[3];
  sum_partials<3>(part, nb, sh, v);
  if (threadIdx.x == 0) {
    double bb = v[0];
    if (singular) {
      bb = v[0] - v[2] * v[2] / c->nwet;
      if (bb < 0.0) bb = 0.0;
    }
    c->bb = bb;
    c->mean_r = singular ? v[1] / c->nwet : 0.0;
  }
}

// ---------------------------------------------------------------------------
// down pass on a level: [CG update] -> pre-smooth -> residual -> restrict
// ---------------------------------------------------------------------------
template <class T>
struct DownArgs {
  LevelDev F, C;
  const T* r_in;
  T* r_out;     // CG: updated residual (ping-pong with r_in)
  const T* q;   // CG
  const T* p;   // CG
  T* x;         // CG iterate
  T* zpre;      // pre-smoothed z on the core
  T* rc;        // restricted residual (coarse level)
  Ctl* ctl;
  const Ctl* outer;
  const double* part_in;  // CG, k > 0: partials of k_pupdate (pq, sum q)
  double* part_out;       // CG: partial ||r||^2
  int nb_in;
  int first;    // CG: iteration 0 (no update; singular projection)
  int iter;
  int inner;    // inner solve of a refinement: first update writes x
  int singular;
};

template <class T, int NU, bool CG, bool COARSE, int FC>
__global__ __launch_bounds__(BS) void k_down(DownArgs<T> a) {
  constexpr int CT = FC / 2, R = 2 * NU + 1, W = FC + 2 * R;
  __shared__ T sr[W * W];
  __shared__ T sz[W * W];
  __shared__ int8_t sm[W * W];
  __shared__ double sh[BS];
  if (stop(a.ctl, a.outer)) return;
  const LevelDev& F = a.F;
  const int t = threadIdx.x;
  const int cy0 = blockIdx.y * CT, cx0 = blockIdx.x * CT;
  const int fy0 = 2 * cy0, fx0 = 2 * cx0;
  T alpha = T(0), shift = T(0), mq = T(0);
  bool zero_x = false;
  if (CG) {
    Ctl* c = a.ctl;
    if (a.first) {
      shift = T(c->mean_r);
      zero_x = c->bb == 0.0;
    } else {
      double v[2];
      sum_partials<2>(a.part_in, a.nb_in, sh, v);
      double rzk = c->rz[(a.iter + 1) & 1];
      double al = v[0] > 0.0 ? rzk / v[0] : 0.0;
      alpha = T(al);
      mq = T(a.singular ? v[1] / c->nwet : 0.0);
      if (blockIdx.x == 0 && blockIdx.y == 0 && t == 0) {
        c->alpha = al;
        c->mean_q = (double)mq;
      }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && t == 0) {
      c->iters = a.iter;
      if (!a.first && a.inner) c->x_valid = 1;
    }
  }
  // window loads: all issued before any use (no dependent chains)
  {
    constexpr int NLD = (W * W + BS - 1) / BS;
    int64_t gi[NLD];
    bool ok[NLD];
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int k = t + i * BS;
      const int ly = k / W, lx = k - ly * W;
      int y = fy0 - R + ly, x = fx0 - R + lx;
      ok[i] = k < W * W && lw(F, y, x);
      gi[i] = ok[i] ? li(F, y, x) : F.origin;
    }
    int8_t mv[NLD];
    T rv[NLD], qv[NLD];
    const bool useq = CG && !a.first;
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      mv[i] = ok[i] ? F.m[gi[i]] : int8_t(0);
      rv[i] = ok[i] ? a.r_in[gi[i]] : T(0);
      qv[i] = (useq && ok[i]) ? a.q[gi[i]] : T(0);
    }
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int k = t + i * BS;
      if (k >= W * W) continue;
      T v = T(0);
      if (mv[i]) {
        v = rv[i];
        if (CG) v = a.first ? v - shift : v - alpha * (qv[i] - mq);
      }
      sr[k] = v;
      sm[k] = mv[i];
      sz[k] = T(0);
    }
  }
  __syncthreads();

  double rr = 0.0;
  if (CG) {
    constexpr int NC = (FC * FC + BS - 1) / BS;
    int64_t gi[NC];
    bool ok[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
      const int e = t + i * BS;
      const int cy = e / FC, cx = e - cy * FC;
      const int y = fy0 + cy, x = fx0 + cx;
      ok[i] = e < FC * FC && y < F.ny && x < F.nx;
      gi[i] = ok[i] ? li(F, y, x) : F.origin;
    }
    T pv[NC], xv[NC];
    const bool upd = !zero_x && !a.first;
    const bool xold = upd && !(a.inner && a.iter == 1);
#pragma unroll
    for (int i = 0; i < NC; i++) {
      pv[i] = (upd && ok[i]) ? a.p[gi[i]] : T(0);
      xv[i] = (xold && ok[i]) ? a.x[gi[i]] : T(0);
    }
#pragma unroll
    for (int i = 0; i < NC; i++) {
      if (!ok[i]) continue;
      const int e = t + i * BS;
      const int cy = e / FC, cx = e - cy * FC;
      const int k = (cy + R) * W + (cx + R);
      const T rv = sr[k];
      a.r_out[gi[i]] = rv;
      if (sm[k]) {
        if (zero_x) a.x[gi[i]] = T(0);
        else if (upd) a.x[gi[i]] = xold ? xv[i] + alpha * pv[i] : alpha * pv[i];
        rr += (double)rv * (double)rv;
      }
    }
  }

  const T cx = T(F.cx), cy = T(F.cy), id = T(F.inv_d), dd = T(F.d);
  // pre-smoothing: half-sweep h updates colour (h-1)&1 within radius R-h+1
#pragma unroll
  for (int h = 1; h <= 2 * NU; h++) {
    const int rho = R - h + 1, col = (h - 1) & 1;
    const int lo = R - rho, hi = R + FC + rho;
    for (int k = t; k < W * W; k += BS) {
      int ly = k / W, lx = k - ly * W;
      if (ly < lo || ly >= hi || lx < lo || lx >= hi || !sm[k]) continue;
      if (((fy0 + ly + fx0 + lx - 2 * R) & 1) != col) continue;
      if (h == 1)
        sz[k] = sr[k] * id;
      else
        sz[k] = (sr[k] + cx * (sz[k - 1] + sz[k + 1]) +
                 cy * (sz[k - W] + sz[k + W])) * id;
    }
    __syncthreads();
  }
  if (COARSE) {
    // residual within radius 1 (into sr)
    for (int k = t; k < W * W; k += BS) {
      int ly = k / W, lx = k - ly * W;
      if (ly < R - 1 || ly >= R + FC + 1 || lx < R - 1 || lx >= R + FC + 1)
        continue;
      T res = T(0);
      if (sm[k])
        res = sr[k] - (dd * sz[k] - cx * (sz[k - 1] + sz[k + 1]) -
                       cy * (sz[k - W] + sz[k + W]));
      sr[k] = res;
    }
    __syncthreads();
    const LevelDev& C = a.C;
    for (int e = t; e < CT * CT; e += BS) {
      int jj = e / CT, ii = e - jj * CT;
      int J = cy0 + jj, I = cx0 + ii;
      if (J >= C.ny || I >= C.nx) continue;
      int64_t cg = li(C, J, I);
      int k = (R + 2 * jj) * W + (R + 2 * ii);
      T v = sr[k] + T(0.5) * ((sr[k - 1] + sr[k + 1]) + (sr[k - W] + sr[k + W])) +
            T(0.25) * ((sr[k - W - 1] + sr[k - W + 1]) +
                       (sr[k + W - 1] + sr[k + W + 1]));
      a.rc[cg] = C.m[cg] ? T(0.25) * v : T(0);
    }
  }
  for (int e = t; e < FC * FC; e += BS) {
    int cy_ = e / FC, cx_ = e - cy_ * FC;
    int y = fy0 + cy_, x = fx0 + cx_;
    if (y >= F.ny || x >= F.nx) continue;
    a.zpre[li(F, y, x)] = sz[(cy_ + R) * W + (cx_ + R)];
  }
  if (!CG) return;
  double brr = block_sum<BS>(rr, sh);
  if (t == 0) a.part_out[blockIdx.y * gridDim.x + blockIdx.x] = brr;
}

// ---------------------------------------------------------------------------
// up pass: x = zpre + P e_coarse, post-smooth; on level 0 the convergence
// test first, then (r,z), sum z, sum r partials
// ---------------------------------------------------------------------------
template <class T>
struct UpArgs {
  LevelDev F, C;
  const T* r;     // right-hand side of this level (level 0: current residual)
  const T* zpre;
  const T* ec;    // coarse correction
  T* out;         // correction of this level (level 0: z)
  Ctl* ctl;
  const Ctl* outer;
  const double* part_in;  // FINE: ||r||^2 partials of k_down<CG>
  double* part_out;       // FINE: (r,z), sum z, sum r
  int nb_in;
  int bench;              // timing run: no convergence test
};

template <class T, int NU, bool FINE, bool COARSE, int FC>
__global__ __launch_bounds__(BS) void k_up(UpArgs<T> a) {
  constexpr int R = 2 * NU, W = FC + 2 * R;
  constexpr int CW = FC / 2 + 2 * NU + 2;
  __shared__ T sx[W * W];
  __shared__ T sr[W * W];
  __shared__ int8_t sm[W * W];
  __shared__ T sc[COARSE ? CW * CW : 1];
  __shared__ double sh[BS];
  if (stop(a.ctl, a.outer)) return;
  const LevelDev& F = a.F;
  const int t = threadIdx.x;
  if (FINE) {
    double v[1];
    sum_partials<1>(a.part_in, a.nb_in, sh, v);
    Ctl* c = a.ctl;
    double rel = c->bb > 0.0 ? sqrt(v[0] / c->bb) : 0.0;
    bool conv = !a.bench && (c->bb == 0.0 || rel <= c->rtol || !(rel == rel));
    if (blockIdx.x == 0 && blockIdx.y == 0 && t == 0) {
      c->rr = v[0];
      c->relres = rel;
      if (conv) c->done = 1;
    }
    if (conv) return;
  }
  const int fy0 = blockIdx.y * FC, fx0 = blockIdx.x * FC;
  const int cyb = fy0 / 2 - NU - 1, cxb = fx0 / 2 - NU - 1;
  {
    constexpr int NLD = (W * W + BS - 1) / BS;
    constexpr int NCL = COARSE ? (CW * CW + BS - 1) / BS : 0;
    int64_t gi[NLD];
    bool ok[NLD];
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int k = t + i * BS;
      const int ly = k / W, lx = k - ly * W;
      int y = fy0 - R + ly, x = fx0 - R + lx;
      ok[i] = k < W * W && lw(F, y, x);
      gi[i] = ok[i] ? li(F, y, x) : F.origin;
    }
    int64_t ci[NCL > 0 ? NCL : 1];
    bool cok[NCL > 0 ? NCL : 1];
#pragma unroll
    for (int i = 0; i < NCL; i++) {
      const LevelDev& C = a.C;
      const int k = t + i * BS;
      const int ly = k / CW, lx = k - ly * CW;
      int y = cyb + ly, x = cxb + lx;
      cok[i] = k < CW * CW && lw(C, y, x);
      ci[i] = cok[i] ? li(C, y, x) : C.origin;
    }
    int8_t mv[NLD];
    T rv[NLD], zv[NLD], cv[NCL > 0 ? NCL : 1];
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      mv[i] = ok[i] ? F.m[gi[i]] : int8_t(0);
      rv[i] = ok[i] ? a.r[gi[i]] : T(0);
      zv[i] = ok[i] ? a.zpre[gi[i]] : T(0);
    }
#pragma unroll
    for (int i = 0; i < NCL; i++) cv[i] = cok[i] ? a.ec[ci[i]] : T(0);
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int k = t + i * BS;
      if (k >= W * W) continue;
      sr[k] = mv[i] ? rv[i] : T(0);
      sx[k] = mv[i] ? zv[i] : T(0);
      sm[k] = mv[i];
    }
#pragma unroll
    for (int i = 0; i < NCL; i++) {
      const int k = t + i * BS;
      if (k < CW * CW) sc[k] = cv[i];
    }
  }
  __syncthreads();
  if (COARSE) {
    for (int k = t; k < W * W; k += BS) {
      if (!sm[k]) continue;
      int ly = k / W, lx = k - ly * W;
      int gy = fy0 - R + ly, gx = fx0 - R + lx;
      int J = (gy >> 1) - cyb, I = (gx >> 1) - cxb;
      const T* c0 = sc + J * CW + I;
      T v;
      bool oy = gy & 1, ox = gx & 1;
      if (!oy && !ox)
        v = c0[0];
      else if (!oy)
        v = T(0.5) * (c0[0] + c0[1]);
      else if (!ox)
        v = T(0.5) * (c0[0] + c0[CW]);
      else
        v = T(0.25) * ((c0[0] + c0[1]) + (c0[CW] + c0[CW + 1]));
      sx[k] = sx[k] + v;
    }
    __syncthreads();
  }
  const T cx = T(F.cx), cy = T(F.cy), id = T(F.inv_d);
#pragma unroll
  for (int h = 1; h <= 2 * NU; h++) {
    const int rho = R - h, col = h & 1;
    const int lo = R - rho, hi = R + FC + rho;
    for (int k = t; k < W * W; k += BS) {
      int ly = k / W, lx = k - ly * W;
      if (ly < lo || ly >= hi || lx < lo || lx >= hi || !sm[k]) continue;
      if (((fy0 + ly + fx0 + lx - 2 * R) & 1) != col) continue;
      sx[k] = (sr[k] + cx * (sx[k - 1] + sx[k + 1]) +
               cy * (sx[k - W] + sx[k + W])) * id;
    }
    __syncthreads();
  }
  double rz = 0.0, zs = 0.0, rs = 0.0;
  for (int e = t; e < FC * FC; e += BS) {
    int cy_ = e / FC, cx_ = e - cy_ * FC;
    int y = fy0 + cy_, x = fx0 + cx_;
    if (y >= F.ny || x >= F.nx) continue;
    int k = (cy_ + R) * W + (cx_ + R);
    T v = sx[k];
    a.out[li(F, y, x)] = v;
    if (FINE && sm[k]) {
      rz += (double)sr[k] * (double)v;
      zs += (double)v;
      rs += (double)sr[k];
    }
  }
  if (!FINE) return;
  double b0 = block_sum<BS>(rz, sh);
  double b1 = block_sum<BS>(zs, sh);
  double b2 = block_sum<BS>(rs, sh);
  const int nb = gridDim.x * gridDim.y, bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (t == 0) {
    a.part_out[bid] = b0;
    a.part_out[nb + bid] = b1;
    a.part_out[2 * nb + bid] = b2;
  }
}

// ---------------------------------------------------------------------------
// p = (z - mean z) + beta p_old ;  q = K p ;  partial (p, q), sum q
// ---------------------------------------------------------------------------
template <class T>
struct PArgs {
  LevelDev L;
  const T* z;
  const T* p_old;
  T* p;
  T* q;
  Ctl* ctl;
  const Ctl* outer;
  const double* part_in;  // (r,z), sum z, sum r of k_up<FINE>
  double* part_out;       // (p,q), sum q
  int nb_in;
  int first;
  int iter;
  int singular;
  int bench;
};

template <class T>
__global__ __launch_bounds__(BS) void k_pupdate(PArgs<T> a) {
  __shared__ double sh[BS];
  if (stop(a.ctl, a.outer)) return;
  Ctl* c = a.ctl;
  double v[3];
  sum_partials<3>(a.part_in, a.nb_in, sh, v);
  const double mzd = a.singular ? v[1] / c->nwet : 0.0;
  // singular operator: the search direction uses z - mean(z), so the scalar
  // is (r, z - mz) = (r, z) - mz * sum(r)
  const double rzk = a.singular ? v[0] - mzd * v[2] : v[0];
  const double rzo = c->rz[(a.iter + 1) & 1];
  const double bet = a.first ? 0.0 : (rzo != 0.0 ? rzk / rzo : 0.0);
  const bool brk = !a.bench && !(rzk > 0.0);  // breakdown / exact solution
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (bid == 0 && threadIdx.x == 0) {
    c->rz[a.iter & 1] = rzk;
    c->beta = bet;
    c->mean_z = mzd;
    if (brk) c->done = 1;
  }
  if (brk) return;
  const LevelDev& L = a.L;
  const T beta = T(bet), mz = T(mzd);
  int y0, x0;
  tile_origin<FC0>(y0, x0);
  // p on the tile and its one-point ring, in LDS
  constexpr int W = FC0 + 2, NLD = (W * W + BS - 1) / BS;
  __shared__ T sp[W * W];
  {
    int64_t gi[NLD];
    bool ok[NLD];
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int k = threadIdx.x + i * BS;
      const int ly = k / W, lx = k - ly * W;
      int y = y0 - 1 + ly, x = x0 - 1 + lx;
      ok[i] = k < W * W && lw(L, y, x);
      gi[i] = ok[i] ? li(L, y, x) : L.origin;
    }
    int8_t mv[NLD];
    T zv[NLD], pv[NLD];
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      mv[i] = ok[i] ? L.m[gi[i]] : int8_t(0);
      zv[i] = ok[i] ? a.z[gi[i]] : T(0);
      pv[i] = (ok[i] && !a.first) ? a.p_old[gi[i]] : T(0);
    }
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int k = threadIdx.x + i * BS;
      if (k >= W * W) continue;
      T v = zv[i] - mz;
      if (!a.first) v = v + beta * pv[i];
      sp[k] = mv[i] ? v : T(0);
    }
  }
  __syncthreads();
  double pq = 0.0, qs = 0.0;
  for (int e = threadIdx.x; e < FC0 * FC0; e += BS) {
    const int cy = e / FC0, cx = e - cy * FC0;
    const int y = y0 + cy, x = x0 + cx;
    if (y >= L.ny || x >= L.nx) continue;
    const int64_t g = li(L, y, x);
    const int k = (cy + 1) * W + cx + 1;
    const T pc = sp[k];
    T qv = T(0);
    if (L.m[g]) {
      qv = T(L.d) * pc - T(L.cx) * (sp[k - 1] + sp[k + 1]) -
           T(L.cy) * (sp[k - W] + sp[k + W]);
      pq += (double)pc * (double)qv;
      qs += (double)qv;
    }
    a.p[g] = pc;
    a.q[g] = qv;
  }
  double b0 = block_sum<BS>(pq, sh);
  double b1 = block_sum<BS>(qs, sh);
  const int nb = gridDim.x * gridDim.y;
  if (threadIdx.x == 0) {
    a.part_out[bid] = b0;
    a.part_out[nb + bid] = b1;
  }
}

// one block after the last iteration: final residual and statistics
__global__ __launch_bounds__(BS) void k_finish(Ctl* c, const double* part_rr,
                                               int nb) {
  __shared__ double sh[BS];
  double v[1];
  sum_partials<1>(part_rr, nb, sh, v);
  if (threadIdx.x != 0) return;
  double rel = c->bb > 0.0 ? sqrt(v[0] / c->bb) : 0.0;
  c->rr = v[0];
  c->relres = rel;
  double conv = (c->bb == 0.0 || rel <= c->rtol) ? 1.0 : 0.0;
  if (conv > 0.0) c->done = 1;
  c->stats[0] = (double)c->iters;
  c->stats[1] = rel;
  c->stats[2] = conv;
  c->stats[3] = 0.0;
  c->stats[4] += conv > 0.0 ? 0.0 : 1.0;
  if ((double)c->iters > c->stats[5]) c->stats[5] = (double)c->iters;
  c->stats[6] += 1.0;
}

// ---------------------------------------------------------------------------
// tail: levels Lt .. L-1 of the V-cycle in one workgroup, all in LDS.
//
// The host writes the V-cycle as a schedule of operations (half-sweep,
// residual, restriction, prolongation, mean projection); the kernel runs the
// schedule with one copy of each operation's code, which keeps it small
// enough for the instruction cache.  Levels larger than WAVE_PTS points run
// on the whole workgroup, smaller ones on wavefront 0 alone (LDS operations
// of one wavefront complete in order, so a wave-level barrier suffices).
// ---------------------------------------------------------------------------
constexpr int MAXL = 16;
constexpr int WAVE_PTS = 256;
constexpr int MAXOPS = 512;
// coarsening stops at a level of at most DENSE_PTS points; that level is
// solved exactly with its dense (pseudo-)inverse
constexpr int DENSE_PTS = 64;
constexpr int DENSE_COLS = DENSE_PTS / 4;  // columns per wave, >= 4 waves

enum TailOp { OP_SWEEP = 0, OP_RES = 1, OP_RESTRICT = 2, OP_PROLONG = 3,
              OP_MEAN_R = 4, OP_MEAN_X = 5, OP_DENSE = 6 };

struct TailStep {
  int8_t op, level, col, fromzero, wave;
  int8_t pad[3];
};

// One tail level in LDS: ny x nx points stored inside a one-point ring,
// row pitch P = nx + 2 (point (y, x) at (y + 1) * P + x + 1).  The ring holds
// the periodic images of the edge points, or zeros in closed directions, so
// every stencil reads its neighbours at fixed offsets.
struct TailLev {
  int ny, nx, P, lg;         // lg = log2(nx) when nx is a power of two, else -1
  int xper, yper;
  int off_r, off_x, off_m;   // byte offsets of r, x (T) and the mask (int8)
  int stride;                // the level's mask in device memory:
  int64_t origin;            //   gm[origin + y * stride + x]
  double cx, cy, d, inv_d;
  const int8_t* gm;
};

template <class T>
struct TailArgs {
  TailLev lv[MAXL];          // host copy of the level descriptors
  const TailLev* lvs;        // the same, in device memory
  LevelDev top;    // the first tail level in device memory
  int off_tmp;     // residual scratch (largest level, >= 1024 entries)
  int nl;
  int nsteps;
  const TailStep* steps;  // device schedule
  const T* r_in;
  T* e_out;
  Ctl* ctl;
  const Ctl* outer;
  // coarsest level: x = M r over its nw unknowns (M symmetric, nw x nw,
  // unknown j at point dense_idx[j] = y * nx + x of the level), or nw = 0
  const T* dense;
  const int* dense_idx;
  int dense_n;
};

// point j of a level in row-major order -> (y, x)
__device__ __forceinline__ void yx_of(int j, int nx, int lg, int& y, int& x) {
  if (lg >= 0) {
    y = j >> lg;
    x = j & (nx - 1);
  } else {
    y = j / nx;
    x = j - y * nx;
  }
}

// Store v at point (y, x), index k, of a level array together with its ring
// images: the periodic copy, or a zero on the closed side (the residual
// scratch is shared by all levels, so its ring must not keep a larger
// level's values).  Levels have nx, ny >= 2.
struct RingGeo {  // a level's ring offsets, held in registers
  int nx1, ny1;     // nx - 1, ny - 1
  int ex0, ex1;     // offset of the image of column 0 / column nx - 1
  int ey0, ey1;     // offset of the image of row 0 / row ny - 1
  bool xper, yper;
};

__device__ __forceinline__ RingGeo ring_geo(const TailLev& L) {
  RingGeo g;
  g.nx1 = L.nx - 1;
  g.ny1 = L.ny - 1;
  g.xper = L.xper != 0;
  g.yper = L.yper != 0;
  g.ex0 = g.xper ? L.nx : -1;
  g.ex1 = g.xper ? -L.nx : 1;
  g.ey0 = g.yper ? L.ny * L.P : -L.P;
  g.ey1 = g.yper ? -L.ny * L.P : L.P;
  return g;
}

template <class T>
__device__ __forceinline__ void put(T* a, const RingGeo& g, int k, int y,
                                    int x, T v) {
  a[k] = v;
  const int ex = x == 0 ? g.ex0 : (x == g.nx1 ? g.ex1 : 0);
  const int ey = y == 0 ? g.ey0 : (y == g.ny1 ? g.ey1 : 0);
  if (ex | ey) {
    const T vx = g.xper ? v : T(0), vy = g.yper ? v : T(0);
    const T vc = (g.xper && g.yper) ? v : T(0);
    if (ex) a[k + ex] = vx;
    if (ey) a[k + ey] = vy;
    if (ex && ey) a[k + ex + ey] = vc;
  }
}

__device__ __forceinline__ double lane_bcast(double v, int j) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, j);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ float lane_bcast(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---- tail operations, specialised on LG = log2(nx) of the level (1..6;
// 0: any nx, run-time geometry) so that the row pitch and the neighbour
// offsets are compile-time constants on power-of-two levels ----

template <int LG>
__device__ __forceinline__ void pt_of(int j, int nx, int& y, int& x) {
  if (LG > 0) {
    y = j >> LG;
    x = j & ((1 << LG) - 1);
  } else {
    y = j / nx;
    x = j - y * nx;
  }
}

template <class T, int LG>
__device__ __forceinline__ void t_sweep(const TailLev& L, const RingGeo& G,
                                        unsigned char* smem, int col,
                                        bool fromzero, int tid, int NT) {
  const int nx = LG > 0 ? (1 << LG) : L.nx;
  const int P = nx + 2;
  const int8_t* m = reinterpret_cast<const int8_t*>(smem + L.off_m);
  const T* r = reinterpret_cast<const T*>(smem + L.off_r);
  T* x = reinterpret_cast<T*>(smem + L.off_x);
  const T cx = T(L.cx), cy = T(L.cy), id = T(L.inv_d);
  const bool even = !(nx & 1);
  const int hx = nx >> 1;
  const int cnt = even ? L.ny * hx : L.ny * nx;
  for (int j = tid; j < cnt; j += NT) {
    int y, xx;
    if (even) {
      int i2;
      if (LG > 0) {
        y = j >> (LG - 1);
        i2 = j & (hx - 1);
      } else {
        y = j / hx;
        i2 = j - y * hx;
      }
      xx = 2 * i2 + ((col ^ y) & 1);
    } else {
      y = j / nx;
      xx = j - y * nx;
      if (((y + xx) & 1) != col) continue;
    }
    const int k = (y + 1) * P + xx + 1;
    const bool wet = m[k] != 0;
    T v;
    if (fromzero) {
      v = r[k] * id;
    } else {
      const T sx = x[k - 1] + x[k + 1];
      const T sy = x[k - P] + x[k + P];
      v = (r[k] + cx * sx + cy * sy) * id;
    }
    if (wet) put(x, G, k, y, xx, v);
  }
}

template <class T, int LG>
__device__ __forceinline__ void t_res(const TailLev& L, const RingGeo& G,
                                      unsigned char* smem, T* tmp, int tid,
                                      int NT) {
  const int nx = LG > 0 ? (1 << LG) : L.nx;
  const int P = nx + 2;
  const int8_t* m = reinterpret_cast<const int8_t*>(smem + L.off_m);
  const T* r = reinterpret_cast<const T*>(smem + L.off_r);
  const T* x = reinterpret_cast<const T*>(smem + L.off_x);
  const T cx = T(L.cx), cy = T(L.cy), dd = T(L.d);
  const int n = L.ny * nx;
  for (int j = tid; j < n; j += NT) {
    int y, xx;
    pt_of<LG>(j, nx, y, xx);
    const int k = (y + 1) * P + xx + 1;
    const bool wet = m[k] != 0;
    const T sx = x[k - 1] + x[k + 1];
    const T sy = x[k - P] + x[k + P];
    const T res = r[k] - (dd * x[k] - cx * sx - cy * sy);
    put(tmp, G, k, y, xx, wet ? res : T(0));
  }
}

// full weighting of tmp (fine level L) onto the rhs of level C = L / 2
template <class T, int LG>
__device__ __forceinline__ void t_restrict(const TailLev& L, const TailLev& C,
                                           unsigned char* smem, const T* tmp,
                                           int tid, int NT) {
  const int nx = LG > 0 ? (1 << LG) : L.nx;
  const int P = nx + 2;
  const int cnx = LG > 0 ? (1 << (LG - 1)) : C.nx;
  const int CP = cnx + 2;
  const int8_t* mc = reinterpret_cast<const int8_t*>(smem + C.off_m);
  T* rc = reinterpret_cast<T*>(smem + C.off_r);
  const int nc = C.ny * cnx;
  for (int j = tid; j < nc; j += NT) {
    int J, I;
    pt_of<(LG > 0 ? LG - 1 : 0)>(j, cnx, J, I);
    const int kc = (J + 1) * CP + I + 1;
    const int kf = (2 * J + 1) * P + 2 * I + 1;
    const bool wet = mc[kc] != 0;
    const T v = tmp[kf] +
                T(0.5) * ((tmp[kf - 1] + tmp[kf + 1]) + (tmp[kf - P] + tmp[kf + P])) +
                T(0.25) * ((tmp[kf - P - 1] + tmp[kf - P + 1]) +
                           (tmp[kf + P - 1] + tmp[kf + P + 1]));
    rc[kc] = wet ? T(0.25) * v : T(0);
  }
}

// bilinear interpolation of level C = L / 2, added to x of level L
template <class T, int LG>
__device__ __forceinline__ void t_prolong(const TailLev& L, const TailLev& C,
                                          const RingGeo& G, unsigned char* smem,
                                          int tid, int NT) {
  const int nx = LG > 0 ? (1 << LG) : L.nx;
  const int P = nx + 2;
  const int cnx = LG > 0 ? (1 << (LG - 1)) : C.nx;
  const int CP = cnx + 2;
  const int8_t* m = reinterpret_cast<const int8_t*>(smem + L.off_m);
  T* x = reinterpret_cast<T*>(smem + L.off_x);
  const T* xc = reinterpret_cast<const T*>(smem + C.off_x);
  const int n = L.ny * nx;
  for (int j = tid; j < n; j += NT) {
    int y, xx;
    pt_of<LG>(j, nx, y, xx);
    const int k = (y + 1) * P + xx + 1;
    const bool wet = m[k] != 0;
    const int kc = ((y >> 1) + 1) * CP + (xx >> 1) + 1;
    const bool oy = y & 1, ox = xx & 1;
    const T c00 = xc[kc], c01 = xc[kc + 1];
    const T c10 = xc[kc + CP], c11 = xc[kc + CP + 1];
    // along x on rows J and J + 1, then along y (the halvings are exact,
    // so odd-odd points get ((c00 + c01) + (c10 + c11)) / 4)
    const T a0 = ox ? T(0.5) * (c00 + c01) : c00;
    const T a1 = ox ? T(0.5) * (c10 + c11) : c10;
    const T v = oy ? T(0.5) * (a0 + a1) : a0;
    const T xv = x[k];
    if (wet) put(x, G, k, y, xx, xv + v);
  }
}

#define F2D_TAIL_LG(lg, CALL) \
  switch (lg) {               \
    case 6: CALL(6); break;   \
    case 5: CALL(5); break;   \
    case 4: CALL(4); break;   \
    case 3: CALL(3); break;   \
    case 2: CALL(2); break;   \
    case 1: CALL(1); break;   \
    default: CALL(0); break;  \
  }

template <class T>
__global__ __launch_bounds__(1024) void k_tail(TailArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (stop(a.ctl, a.outer)) return;
  // level descriptors and the schedule are read-only device arrays: their
  // uniform reads go through the scalar cache into scalar registers
  const TailLev* __restrict__ lvs = a.lvs;
  const TailStep* __restrict__ steps = a.steps;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int wave = tid >> 6, lane = tid & 63;
  // zero every level (ring included), then fill masks and the input
  for (int l = 0; l < a.nl; l++) {
    const TailLev& L = lvs[l];
    const int np = (L.ny + 2) * L.P;
    int8_t* m = reinterpret_cast<int8_t*>(smem + L.off_m);
    T* x = reinterpret_cast<T*>(smem + L.off_x);
    for (int k = tid; k < np; k += nth) {
      m[k] = 0;
      x[k] = T(0);
    }
  }
  __syncthreads();
  for (int l = 0; l < a.nl; l++) {
    const TailLev& L = lvs[l];
    const int n = L.ny * L.nx;
    int8_t* m = reinterpret_cast<int8_t*>(smem + L.off_m);
    T* r = reinterpret_cast<T*>(smem + L.off_r);
    for (int j = tid; j < n; j += nth) {
      int y, xx;
      yx_of(j, L.nx, L.lg, y, xx);
      const int k = (y + 1) * L.P + xx + 1;
      const int64_t g = L.origin + (int64_t)y * L.stride + xx;
      const int8_t mv = L.gm[g];
      m[k] = mv;
      if (l == 0) r[k] = mv ? a.r_in[g] : T(0);
    }
  }
  T* tmp = reinterpret_cast<T*>(smem + a.off_tmp);
  const int nsteps = a.nsteps;
  // every wave walks the schedule: a block-mode step, or a wave-mode step
  // that follows a block-mode one, starts with a workgroup barrier (the same
  // decision in every wave); consecutive wave-mode steps are separated by a
  // wave-level barrier in wavefront 0 while the others skip them
  int prev_wave = 0;
#pragma unroll 1
  for (int i = 0; i < nsteps; i++) {
    const TailStep st = steps[i];
    if (!st.wave || !prev_wave) __syncthreads();
    prev_wave = st.wave;
    int NT;
    if (st.wave) {
      if (wave != 0) continue;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      NT = 64;
    } else {
      NT = nth;
    }
    const TailLev& L = lvs[st.level];
    const RingGeo G = ring_geo(L);
    if (st.op == OP_SWEEP) {
#define F2D_CALL(v) t_sweep<T, v>(L, G, smem, st.col, st.fromzero, tid, NT)
      F2D_TAIL_LG(L.lg, F2D_CALL)
#undef F2D_CALL
    } else if (st.op == OP_RES) {
#define F2D_CALL(v) t_res<T, v>(L, G, smem, tmp, tid, NT)
      F2D_TAIL_LG(L.lg, F2D_CALL)
#undef F2D_CALL
    } else if (st.op == OP_RESTRICT) {
      const TailLev& C = lvs[st.level + 1];
#define F2D_CALL(v) t_restrict<T, v>(L, C, smem, tmp, tid, NT)
      F2D_TAIL_LG(L.lg, F2D_CALL)
#undef F2D_CALL
    } else if (st.op == OP_PROLONG) {
      const TailLev& C = lvs[st.level + 1];
#define F2D_CALL(v) t_prolong<T, v>(L, C, G, smem, tid, NT)
      F2D_TAIL_LG(L.lg, F2D_CALL)
#undef F2D_CALL
    } else if (st.op == OP_DENSE) {
      // block mode: wave w forms the partial products of columns
      // [w cpw, (w + 1) cpw), lane i owns unknown i; wave 0 adds the partials
      const int P = L.P;
      const T* r = reinterpret_cast<const T*>(smem + L.off_r);
      T* x = reinterpret_cast<T*>(smem + L.off_x);
      const int nw = a.dense_n, nwv = nth >> 6;
      const int cpw = (nw + nwv - 1) / nwv;
      const int j0 = wave * cpw;
      const bool act = lane < nw;
      int y = 0, xx = 0;
      if (act) yx_of(a.dense_idx[lane], L.nx, L.lg, y, xx);
      const int k = (y + 1) * P + xx + 1;
      const T ri = act ? r[k] : T(0);
      T mv[DENSE_COLS];
#pragma unroll
      for (int u = 0; u < DENSE_COLS; u++) {
        const int j = j0 + u;
        mv[u] = (u < cpw && j < nw && act) ? a.dense[j * nw + lane] : T(0);
      }
      T part = T(0);
#pragma unroll
      for (int u = 0; u < DENSE_COLS; u++) {
        const int j = j0 + u;
        if (u < cpw && j < nw) part += mv[u] * lane_bcast(ri, j);
      }
      tmp[wave * 64 + lane] = part;
      __syncthreads();
      if (wave == 0 && act) {
        T s = T(0);
        for (int w = 0; w < nwv; w++) s += tmp[w * 64 + lane];
        put(x, G, k, y, xx, s);
      }
    } else {  // OP_MEAN_R / OP_MEAN_X: zero-mean projection (wave mode)
      const int P = L.P, n = L.ny * L.nx;
      const int8_t* m = reinterpret_cast<const int8_t*>(smem + L.off_m);
      const bool isr = st.op == OP_MEAN_R;
      T* v = reinterpret_cast<T*>(smem + (isr ? L.off_r : L.off_x));
      double sm = 0.0, c = 0.0;
      for (int j = tid; j < n; j += 64) {
        int y, xx;
        yx_of(j, L.nx, L.lg, y, xx);
        const int k = (y + 1) * P + xx + 1;
        if (m[k]) {
          sm += (double)v[k];
          c += 1.0;
        }
      }
      sm = wave_sum(sm);
      c = wave_sum(c);
      const double mean = c > 0 ? sm / c : 0.0;
      for (int j = tid; j < n; j += 64) {
        int y, xx;
        yx_of(j, L.nx, L.lg, y, xx);
        const int k = (y + 1) * P + xx + 1;
        if (m[k]) {
          const T nv = v[k] - T(mean);
          if (isr) v[k] = nv;
          else put(v, G, k, y, xx, nv);
        }
      }
    }
  }
  __syncthreads();
  {
    const TailLev& L = lvs[0];
    const T* x = reinterpret_cast<const T*>(smem + L.off_x);
    const int n = L.ny * L.nx;
    for (int j = tid; j < n; j += nth) {
      int y, xx;
      yx_of(j, L.nx, L.lg, y, xx);
      a.e_out[li(a.top, y, xx)] = x[(y + 1) * L.P + xx + 1];
    }
  }
}

// The tail's schedule: pre-smoothing (red, black)^nu, residual, restriction
// down to the coarsest level; there either the dense solve or
// (red, black)^S (black, red)^S (with the zero-mean projection of the
// singular operator); prolongation and post-smoothing (black, red)^nu back up.
static std::vector<TailStep> tail_schedule(const std::vector<int>& npts,
                                           int nu, int coarse_sweeps,
                                           bool singular, bool dense) {
  std::vector<TailStep> v;
  const int nl = (int)npts.size();
  auto add = [&](int op, int l, int col, int fz, int wmode) {
    TailStep s{};
    s.op = (int8_t)op;
    s.level = (int8_t)l;
    s.col = (int8_t)col;
    s.fromzero = (int8_t)fz;
    s.wave = (int8_t)wmode;
    v.push_back(s);
  };
  for (int l = 0; l + 1 < nl; l++) {
    const int w = npts[l] <= WAVE_PTS;
    for (int h = 1; h <= 2 * nu; h++) add(OP_SWEEP, l, (h - 1) & 1, h == 1, w);
    add(OP_RES, l, 0, 0, w);
    add(OP_RESTRICT, l, 0, 0, w);
  }
  if (dense) {
    // the pseudo-inverse annihilates constants only up to rounding in FP32:
    // the explicit projections keep the correction exactly zero-mean
    if (singular) add(OP_MEAN_R, nl - 1, 0, 0, 1);
    add(OP_DENSE, nl - 1, 0, 0, 0);
    if (singular) add(OP_MEAN_X, nl - 1, 0, 0, 1);
  } else {
    const int l = nl - 1;
    const int w = npts[l] <= WAVE_PTS;
    if (singular) add(OP_MEAN_R, l, 0, 0, 1);
    for (int s = 0; s < coarse_sweeps; s++) {
      add(OP_SWEEP, l, 0, s == 0, w);
      add(OP_SWEEP, l, 1, 0, w);
    }
    for (int s = 0; s < coarse_sweeps; s++) {
      add(OP_SWEEP, l, 1, 0, w);
      add(OP_SWEEP, l, 0, 0, w);
    }
    if (singular) add(OP_MEAN_X, l, 0, 0, 1);
  }
  for (int l = nl - 2; l >= 0; l--) {
    const int w = npts[l] <= WAVE_PTS;
    add(OP_PROLONG, l, 0, 0, w);
    for (int h = 1; h <= 2 * nu; h++) add(OP_SWEEP, l, h & 1, 0, w);
  }
  return v;
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement around an FP32 solve
// ---------------------------------------------------------------------------
template <class TB>
__device__ __forceinline__ double g64_at(const TB* b, int64_t k, int64_t P,
                                         int gmode) {
  if (gmode == G_FROM_B) return -(double)b[k];
  double s = 0.25 * ((((double)b[k - P - 1] + (double)b[k - 1]) +
                      (double)b[k - P]) + (double)b[k]);
  return -s;
}

// partial sums of g over the unknowns, in FP64 (the singular operator's
// right-hand side is projected before any FP32 rounding)
template <class TB>
__global__ __launch_bounds__(BS) void k_gsum(LevelDev L, const TB* b,
                                             int gmode, double* part) {
  __shared__ double sh[BS];
  int y0, x0;
  tile_origin<FC0>(y0, x0);
  double s = 0.0;
  for (int e = threadIdx.x; e < FC0 * FC0; e += BS) {
    int y = y0 + e / FC0, x = x0 + e % FC0;
    if (y >= L.ny || x >= L.nx) continue;
    int64_t k = li(L, y, x);
    if (L.m[k]) s += g64_at(b, k, L.stride, gmode);
  }
  double bs = block_sum<BS>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = bs;
}

// one block: out = (sum of partials) / n
__global__ __launch_bounds__(BS) void k_mean_fin(const double* part, int nb,
                                                 double n, double* out) {
  __shared__ double sh[BS];
  double v[1];
  sum_partials<1>(part, nb, sh, v);
  if (threadIdx.x == 0) *out = v[0] / n;
}

template <class TB>
struct ResArgs {
  LevelDev L;
  const TB* b;      // b or omega (padded)
  double* x64a;     // FP64 iterate buffers; outer->xcur selects the current
  double* x64b;     // one, the kernel writes the other
  const float* e;   // inner FP32 correction
  float* r32;       // inner right-hand side (= its initial residual)
  const Ctl* inner;
  const Ctl* outer;
  double* part;
  int gmode, singular, add_e;
  const double* gmean;  // singular operator: mean of g over the unknowns
};

template <class TB>
__device__ __forceinline__ double res_x(const ResArgs<TB>& a,
                                        const double* xin, int y, int x,
                                        bool addv) {
  if (!lw(a.L, y, x)) return 0.0;
  int64_t k = li(a.L, y, x);
  if (!a.L.m[k]) return 0.0;
  double v = xin[k];
  if (addv) v += (double)a.e[k];
  return v;
}

// r = g - K (x + e) in FP64; x + e stored, r rounded to FP32 for the inner
// solve; partials sum g^2, sum r^2, sum r, sum g
template <class TB>
__global__ __launch_bounds__(BS) void k_refine_res(ResArgs<TB> a) {
  __shared__ double sh[BS];
  if (a.outer->odone) return;
  const LevelDev& L = a.L;
  const bool addv = a.add_e && a.inner->x_valid;
  const int cur = a.outer->xcur;
  const double* xin = cur ? a.x64b : a.x64a;
  double* xout = cur ? a.x64a : a.x64b;
  const double gm = a.singular ? *a.gmean : 0.0;
  int y0, x0;
  tile_origin<FC0>(y0, x0);
  double gg = 0.0, rr = 0.0, rs = 0.0, gs = 0.0;
  for (int e = threadIdx.x; e < FC0 * FC0; e += BS) {
    int y = y0 + e / FC0, x = x0 + e % FC0;
    if (y >= L.ny || x >= L.nx) continue;
    int64_t k = li(L, y, x);
    float r32 = 0.f;
    if (L.m[k]) {
      double g = g64_at(a.b, k, L.stride, a.gmode) - gm;
      double xc = res_x(a, xin, y, x, addv);
      double w = res_x(a, xin, y, x - 1, addv), ee = res_x(a, xin, y, x + 1, addv);
      double s = res_x(a, xin, y - 1, x, addv), n = res_x(a, xin, y + 1, x, addv);
      double r = g - (L.d * xc - L.cx * (w + ee) - L.cy * (s + n));
      xout[k] = xc;
      r32 = (float)r;
      gg += g * g;
      rr += r * r;
      rs += r;
      gs += g;
    }
    a.r32[k] = r32;
  }
  double b0 = block_sum<BS>(gg, sh);
  double b1 = block_sum<BS>(rr, sh);
  double b2 = block_sum<BS>(rs, sh);
  double b3 = block_sum<BS>(gs, sh);
  const int nb = gridDim.x * gridDim.y, bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (threadIdx.x == 0) {
    a.part[bid] = b0;
    a.part[nb + bid] = b1;
    a.part[2 * nb + bid] = b2;
    a.part[3 * nb + bid] = b3;
  }
}

// one block after each k_refine_res: outer test, buffer flip, reset of the
// inner control block (its right-hand side is the new residual)
__global__ __launch_bounds__(BS) void k_refine_check(Ctl* o, Ctl* c,
                                                     const double* part, int nb,
                                                     int singular, int last) {
  __shared__ double sh[BS];
  if (o->odone) return;
  double v[4];
  sum_partials<4>(part, nb, sh, v);
  if (threadIdx.x != 0) return;
  double tg = v[0], tr = v[1], ts = v[2], tgs = v[3];
  if (singular) {  // norms of the zero-mean projections
    tg = tg - tgs * tgs / c->nwet;
    tr = tr - ts * ts / c->nwet;
    if (tg < 0.0) tg = 0.0;
    if (tr < 0.0) tr = 0.0;
  }
  o->obb = tg;
  o->orr = tr;
  double rel = tg > 0.0 ? sqrt(tr / tg) : 0.0;
  o->orelres = rel;
  o->orefine += 1;
  if (o->orefine > 1) o->inner_total += c->iters;
  o->xcur ^= 1;
  bool conv = tg == 0.0 || rel <= o->ortol || !(rel == rel);
  if (conv || last) o->odone = 1;
  c->bb = tr;
  c->rtol = o->inner_rtol;
  c->mean_r = singular ? ts / c->nwet : 0.0;
  c->done = (conv || last) ? 1 : 0;
  c->iters = 0;
  c->x_valid = 0;
  c->rz[0] = c->rz[1] = 0.0;
}

// write the refined solution (minus `shift`, the gauge of the singular
// operator) to the caller's array at the unknowns
template <class TX>
__global__ void k_refine_out(LevelDev L, const double* x64a,
                             const double* x64b, const Ctl* outer,
                             const double* shift, TX* x) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)L.ny * L.nx) return;
  int y = (int)(p / L.nx), xx = (int)(p - (int64_t)y * L.nx);
  int64_t k = li(L, y, xx);
  const double* x64 = outer->xcur ? x64b : x64a;
  double sh = shift ? *shift : 0.0;
  if (L.m[k]) x[k] = TX(x64[k] - sh);
}

__global__ void k_reset_outer(Ctl* o, double nwet, double rtol,
                              double inner_rtol) {
  if (threadIdx.x != 0) return;
  o->odone = 0;
  o->orefine = 0;
  o->inner_total = 0;
  o->obb = o->orr = 0.0;
  o->orelres = 1.0;
  o->nwet = nwet;
  o->ortol = rtol;
  o->inner_rtol = inner_rtol;
}

__global__ void k_finish_refine(Ctl* c, const Ctl* o) {
  if (threadIdx.x != 0) return;
  double conv = (o->orelres <= o->ortol || o->obb == 0.0) ? 1.0 : 0.0;
  c->stats[0] = (double)o->inner_total;
  c->stats[1] = o->orelres;
  c->stats[2] = conv;
  c->stats[3] = (double)o->orefine;
  c->stats[4] += conv > 0.0 ? 0.0 : 1.0;
  if ((double)o->orefine > c->stats[5]) c->stats[5] = (double)o->orefine;
  c->stats[6] += 1.0;
}

// partial sums of the current FP64 iterate (singular gauge)
__global__ __launch_bounds__(BS) void k_sum64(LevelDev L, const double* x64a,
                                              const double* x64b,
                                              const Ctl* outer, double* part) {
  __shared__ double sh[BS];
  const double* x = outer->xcur ? x64b : x64a;
  int y0, x0;
  tile_origin<FC0>(y0, x0);
  double s = 0.0;
  for (int e = threadIdx.x; e < FC0 * FC0; e += BS) {
    int y = y0 + e / FC0, xx = x0 + e % FC0;
    if (y >= L.ny || xx >= L.nx) continue;
    int64_t k = li(L, y, xx);
    if (L.m[k]) s += x[k];
  }
  double b = block_sum<BS>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = b;
}

template <class TX>
__global__ void k_refine_seed(LevelDev L, const TX* x, double* x64) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)L.ny * L.nx) return;
  int y = (int)(p / L.nx), xx = (int)(p - (int64_t)y * L.nx);
  int64_t k = li(L, y, xx);
  double v = 0.0;
  if (L.m[k]) {
    v = (double)x[k];
    if (!isfinite(v)) v = 0.0;
  }
  x64[k] = v;
}

// gauge of the singular operator: x -= mean(x) over the unknowns
template <class T>
__global__ __launch_bounds__(BS) void k_sum(LevelDev L, const T* x,
                                            double* part) {
  __shared__ double sh[BS];
  int y0, x0;
  tile_origin<FC0>(y0, x0);
  double s = 0.0;
  for (int e = threadIdx.x; e < FC0 * FC0; e += BS) {
    int y = y0 + e / FC0, xx = x0 + e % FC0;
    if (y >= L.ny || xx >= L.nx) continue;
    int64_t k = li(L, y, xx);
    if (L.m[k]) s += (double)x[k];
  }
  double b = block_sum<BS>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = b;
}

template <class T>
__global__ void k_sub_mean(LevelDev L, T* x, const double* mean) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)L.ny * L.nx) return;
  int y = (int)(p / L.nx), xx = (int)(p - (int64_t)y * L.nx);
  int64_t k = li(L, y, xx);
  if (L.m[k]) x[k] = x[k] - T(*mean);
}

__global__ void k_reset_stats(Ctl* c) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < 8; i++) c->stats[i] = 0.0;
}

// fill the halo of the solution (Poisson2D.solve calls mesh.fill)
template <class T>
__global__ void k_fill_x(T* a, int ny, int nx, int xper, int yper) {
  const int NX = nx + 6, NY = ny + 6;
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)NY * NX) return;
  int j = (int)(k / NX), i = (int)(k - (int64_t)j * NX);
  int js = j, is = i;
  if (xper) is = (i < 3) ? i + nx : (i >= nx + 3 ? i - nx : i);
  if (yper) js = (j < 3) ? j + ny : (j >= ny + 3 ? j - ny : j);
  if (js == j && is == i) return;
  a[k] = a[(int64_t)js * NX + is];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct HostLevel {
  LevelDev d;
  int8_t* m = nullptr;
  void* r = nullptr;   // rhs (l >= 1)
  void* z = nullptr;   // pre-smoothed (all levels)
  void* e = nullptr;   // correction (l >= 1)
  size_t n = 0;        // storage entries
  int fc = 32;         // tile of the multi-block kernels on this level
};

template <class T>
struct Pcg {
  std::vector<HostLevel> lv;
  int L = 0, Lt = 0;
  T* r[2] = {nullptr, nullptr};
  T* p[2] = {nullptr, nullptr};
  T* q = nullptr;
  T* z = nullptr;
  TailArgs<T> tail{};
  size_t tail_lds = 0;
  Ctl* ctl = nullptr;
};

}  // namespace f2d

struct f2d_poisson {
  f2d::LevelDev base{};  // level 0 (padded layout)
  int64_t ny = 0, nx = 0, NYP = 0, NXP = 0;
  int xper = 0, yper = 0;
  double maindiag = 0.0;
  int singular = 0;
  double nwet = 0.0;
  f2d_poisson_opts opts{};
  int8_t* m0 = nullptr;
  f2d::Pcg<double> p64;
  f2d::Pcg<float> p32;
  double* x64[2] = {nullptr, nullptr};
  float* e32 = nullptr;
  double* part = nullptr;    // partial-sum regions
  int64_t region = 0;        // doubles per region
  double* scal = nullptr;    // device scalars
  f2d::Ctl* outer = nullptr;
  f2d::Ctl* bench_ctl = nullptr;  // scratch control block for kernel timing
  int have_x64 = 0;
  int last_iters = 0;
  int nu_coarse = 2, nu_tail = 2;  // smoothing sweeps: multi-block coarse
                                   // levels, single-workgroup tail
  double inner_rtol = 1e-4;
  f2d::FftSolver* fft = nullptr;  // direct spectral solve, when it applies
  f2d::NeumannSolver* neu = nullptr;  // Neumann (cell) operator: Jacobi-PCG
  f2d::CapSolver* cap = nullptr;  // masked operator: spectral + capacitance
  int64_t fft_unmeasured = 0;     // FFT solves made without statistics
  int refine_next = -1;  // FP32 direct solves: refinements (0: none), -1 options
  std::vector<void*> allocs;
  double* reg(int r) { return part + (int64_t)r * region; }
};

namespace f2d {

static void* dalloc(f2d_poisson* h, size_t bytes) {
  void* p = nullptr;
  F2D_HIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
  F2D_HIP(hipMemset(p, 0, bytes < 256 ? 256 : bytes));
  h->allocs.push_back(p);
  return p;
}

static dim3 tiles(const LevelDev& L, int fc) {
  return dim3((unsigned)((L.nx + fc - 1) / fc), (unsigned)((L.ny + fc - 1) / fc));
}

// Dense inverse of a coarse level's operator K (d on the diagonal, -cx / -cy
// for each wet neighbour, wrapped in periodic directions; a neighbour met
// twice through the wrap counts twice) over its wet points, row-major in
// `idx` order.  For the singular operator (all wet, doubly periodic, no
// diagonal shift) the pseudo-inverse (K + u u^T)^-1 - u u^T, u = 1/sqrt(nw):
// it annihilates constants and returns zero-mean corrections.
static void coarse_inverse(const TailLev& L, const std::vector<int8_t>& m,
                           bool singular, std::vector<int>& idx,
                           std::vector<double>& M) {
  const int ny = L.ny, nx = L.nx;
  std::vector<int> pos((size_t)ny * nx, -1);
  idx.clear();
  for (int k = 0; k < ny * nx; k++)
    if (m[k]) {
      pos[k] = (int)idx.size();
      idx.push_back(k);
    }
  const int n = (int)idx.size();
  F2D_REQUIRE(n <= DENSE_PTS, "coarse level too large for the dense solve");
  std::vector<double> A((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) {
    const int y = idx[i] / nx, x = idx[i] % nx;
    A[(size_t)i * n + i] += L.d;
    const int dy[4] = {0, 0, -1, 1}, dx[4] = {-1, 1, 0, 0};
    for (int q = 0; q < 4; q++) {
      int yy = y + dy[q], xx = x + dx[q];
      if (xx < 0 || xx >= nx) {
        if (!L.xper) continue;
        xx = (xx + nx) % nx;
      }
      if (yy < 0 || yy >= ny) {
        if (!L.yper) continue;
        yy = (yy + ny) % ny;
      }
      const int j = pos[yy * nx + xx];
      if (j >= 0) A[(size_t)i * n + j] -= (q < 2 ? L.cx : L.cy);
    }
  }
  const double u2 = singular ? 1.0 / n : 0.0;
  if (singular)
    for (auto& v : A) v += u2;
  // Cholesky A = C C^T, then M = A^-1 column by column
  std::vector<double> C((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++) {
    double s = A[(size_t)j * n + j];
    for (int k = 0; k < j; k++) s -= C[(size_t)j * n + k] * C[(size_t)j * n + k];
    F2D_REQUIRE(s > 0.0, "coarse operator is not positive definite");
    const double cjj = std::sqrt(s);
    C[(size_t)j * n + j] = cjj;
    for (int i = j + 1; i < n; i++) {
      double t = A[(size_t)i * n + j];
      for (int k = 0; k < j; k++) t -= C[(size_t)i * n + k] * C[(size_t)j * n + k];
      C[(size_t)i * n + j] = t / cjj;
    }
  }
  M.assign((size_t)n * n, 0.0);
  std::vector<double> w(n);
  for (int c = 0; c < n; c++) {
    for (int i = 0; i < n; i++) {
      double t = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) t -= C[(size_t)i * n + k] * w[k];
      w[i] = t / C[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
      double t = w[i];
      for (int k = i + 1; k < n; k++) t -= C[(size_t)k * n + i] * w[k];
      w[i] = t / C[(size_t)i * n + i];
    }
    for (int i = 0; i < n; i++) M[(size_t)i * n + c] = w[i] - u2;
  }
  // exact symmetry (the kernel reads M by columns)
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      const double v = 0.5 * (M[(size_t)i * n + j] + M[(size_t)j * n + i]);
      M[(size_t)i * n + j] = M[(size_t)j * n + i] = v;
    }
}

template <class T>
static void build_pcg(f2d_poisson* h, Pcg<T>& P) {
  const int tail_points = h->opts.tail_points > 0 ? h->opts.tail_points : 1024;
  HostLevel l0;
  l0.d = h->base;
  l0.m = h->m0;
  l0.d.m = h->m0;
  l0.n = (size_t)(h->NYP * h->NXP);
  l0.fc = FC0;
  P.lv.push_back(l0);
  std::vector<int8_t> mprev((size_t)h->ny * h->nx);
  {
    std::vector<int8_t> full((size_t)(h->NYP * h->NXP));
    F2D_HIP(hipMemcpy(full.data(), h->m0, full.size(), hipMemcpyDeviceToHost));
    for (int64_t y = 0; y < h->ny; y++)
      for (int64_t x = 0; x < h->nx; x++)
        mprev[y * h->nx + x] = full[(y + 3) * h->NXP + x + 3];
  }
  int ny = (int)h->ny, nx = (int)h->nx;
  double cx = h->base.cx, cy = h->base.cy;
  while ((int)P.lv.size() < MAXL) {
    if ((ny & 1) || (nx & 1) || ny < 4 || nx < 4) break;
    if ((int64_t)ny * nx <= DENSE_PTS && P.lv.size() > 1) break;
    int cny = ny / 2, cnx = nx / 2;
    std::vector<int8_t> mc((size_t)cny * cnx);
    for (int J = 0; J < cny; J++)
      for (int I = 0; I < cnx; I++)
        mc[(size_t)J * cnx + I] = mprev[(size_t)(2 * J) * nx + 2 * I];
    cx *= 0.25;
    cy *= 0.25;
    HostLevel c;
    c.d.ny = cny;
    c.d.nx = cnx;
    c.d.stride = cnx;
    c.d.origin = 0;
    c.d.xper = h->xper;
    c.d.yper = h->yper;
    c.d.cx = cx;
    c.d.cy = cy;
    c.d.d = 2.0 * (cx + cy) + h->maindiag;
    c.d.inv_d = 1.0 / c.d.d;
    c.n = (size_t)cny * cnx;
    c.m = (int8_t*)dalloc(h, c.n);
    F2D_HIP(hipMemcpy(c.m, mc.data(), c.n, hipMemcpyHostToDevice));
    c.d.m = c.m;
    // small levels: smaller tiles, more blocks in flight
    c.fc = ((size_t)((cny + 31) / 32) * ((cnx + 31) / 32) < 512) ? 16 : 32;
    P.lv.push_back(c);
    mprev.swap(mc);
    ny = cny;
    nx = cnx;
  }
  P.L = (int)P.lv.size();
  // LDS of a tail starting at level l: r, x and mask of each level with
  // their one-point ring, and the residual scratch
  auto padded = [&](int k) {
    return (size_t)(P.lv[k].d.ny + 2) * (size_t)(P.lv[k].d.nx + 2);
  };
  auto al16 = [](size_t b) { return (b + 15) / 16 * 16; };
  auto tail_bytes = [&](int l) {
    size_t bytes = 0, nmax = 1024;
    for (int k = l; k < P.L; k++) {
      const size_t n = padded(k);
      bytes += 2 * al16(n * sizeof(T)) + al16(n);
      if (n > nmax) nmax = n;
    }
    return bytes + al16(nmax * sizeof(T));
  };
  P.Lt = P.L;
  for (int l = 1; l < P.L; l++) {
    size_t pts = (size_t)P.lv[l].d.ny * P.lv[l].d.nx;
    if ((int)pts > tail_points) continue;
    if (tail_bytes(l) <= 150 * 1024 && P.L - l <= MAXL) {
      P.Lt = l;
      break;
    }
  }
  for (int l = 0; l < P.L; l++) {
    HostLevel& v = P.lv[l];
    v.z = dalloc(h, v.n * sizeof(T));
    if (l >= 1) {
      v.r = dalloc(h, v.n * sizeof(T));
      v.e = dalloc(h, v.n * sizeof(T));
    }
  }
  size_t n0 = P.lv[0].n;
  for (int k = 0; k < 2; k++) {
    P.r[k] = (T*)dalloc(h, n0 * sizeof(T));
    P.p[k] = (T*)dalloc(h, n0 * sizeof(T));
  }
  P.q = (T*)dalloc(h, n0 * sizeof(T));
  P.z = (T*)dalloc(h, n0 * sizeof(T));
  P.ctl = (Ctl*)dalloc(h, sizeof(Ctl));
  Ctl c0{};
  c0.nwet = h->nwet;
  F2D_HIP(hipMemcpy(P.ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice));
  if (P.Lt < P.L) {
    TailArgs<T>& ta = P.tail;
    ta.nl = P.L - P.Lt;
    int off = 0;
    size_t nmax = 1024;
    for (int k = 0; k < ta.nl; k++) {
      const HostLevel& v = P.lv[P.Lt + k];
      F2D_REQUIRE(v.d.ny >= 2 && v.d.nx >= 2, "tail level narrower than 2 points");
      TailLev& t = ta.lv[k];
      t = TailLev{};
      t.ny = v.d.ny;
      t.nx = v.d.nx;
      t.P = v.d.nx + 2;
      t.lg = -1;
      if ((t.nx & (t.nx - 1)) == 0) {
        t.lg = 0;
        while ((1 << t.lg) < t.nx) t.lg++;
      }
      t.xper = v.d.xper;
      t.yper = v.d.yper;
      t.stride = v.d.stride;
      t.origin = v.d.origin;
      t.gm = v.d.m;
      t.cx = v.d.cx;
      t.cy = v.d.cy;
      t.d = v.d.d;
      t.inv_d = v.d.inv_d;
      const size_t n = padded(P.Lt + k);
      if (n > nmax) nmax = n;
      t.off_r = off;
      off += (int)al16(n * sizeof(T));
      t.off_x = off;
      off += (int)al16(n * sizeof(T));
      t.off_m = off;
      off += (int)al16(n);
    }
    ta.top = P.lv[P.Lt].d;
    ta.off_tmp = off;
    off += (int)al16(nmax * sizeof(T));
    P.tail_lds = off;
    const TailLev& cl = ta.lv[ta.nl - 1];
    int cs = h->opts.coarse_sweeps;
    if (cs <= 0) {
      int mx = cl.ny > cl.nx ? cl.ny : cl.nx;
      cs = mx <= 4 ? 2 : (mx + 1) / 2 + 2;
    }
    std::vector<int> npts;
    for (int k = 0; k < ta.nl; k++) npts.push_back(ta.lv[k].ny * ta.lv[k].nx);
    // the coarsest level's dense inverse, when it is small enough
    const bool dense = P.L > 1 && (int64_t)cl.ny * cl.nx <= DENSE_PTS;
    ta.dense = nullptr;
    ta.dense_idx = nullptr;
    ta.dense_n = 0;
    if (dense) {
      std::vector<int> idx;
      std::vector<double> M;
      coarse_inverse(cl, mprev, h->singular, idx, M);
      const int nw = (int)idx.size();
      if (nw > 0) {
        std::vector<T> Mt(M.begin(), M.end());
        T* dM = (T*)dalloc(h, Mt.size() * sizeof(T));
        int* dI = (int*)dalloc(h, idx.size() * sizeof(int));
        F2D_HIP(hipMemcpy(dM, Mt.data(), Mt.size() * sizeof(T), hipMemcpyHostToDevice));
        F2D_HIP(hipMemcpy(dI, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
        ta.dense = dM;
        ta.dense_idx = dI;
      }
      ta.dense_n = nw;
    }
    std::vector<TailStep> sched =
        tail_schedule(npts, h->nu_tail, cs, h->singular, dense);
    F2D_REQUIRE((int)sched.size() <= MAXOPS, "tail schedule too long");
    TailStep* dsched = (TailStep*)dalloc(h, sched.size() * sizeof(TailStep));
    F2D_HIP(hipMemcpy(dsched, sched.data(), sched.size() * sizeof(TailStep),
                      hipMemcpyHostToDevice));
    ta.steps = dsched;
    TailLev* dlv = (TailLev*)dalloc(h, sizeof(TailLev) * ta.nl);
    F2D_HIP(hipMemcpy(dlv, ta.lv, sizeof(TailLev) * ta.nl, hipMemcpyHostToDevice));
    ta.lvs = dlv;
    ta.nsteps = (int)sched.size();
    ta.r_in = (const T*)P.lv[P.Lt].r;
    ta.e_out = (T*)P.lv[P.Lt].e;
    ta.ctl = P.ctl;
  }
}

template <class T>
static void set_tail_attr(size_t bytes) {
  if (bytes > 64 * 1024)
    F2D_HIP(hipFuncSetAttribute((const void*)k_tail<T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
}

template <class T, int NU, bool CG, bool COARSE>
static void launch_down_fc(const DownArgs<T>& a, int fc, hipStream_t s) {
  if (fc == 16)
    k_down<T, NU, CG, COARSE, 16><<<tiles(a.F, 16), BS, 0, s>>>(a);
  else
    k_down<T, NU, CG, COARSE, 32><<<tiles(a.F, 32), BS, 0, s>>>(a);
}

template <class T, int NU>
static void launch_down_nu(const DownArgs<T>& a, bool cg, bool coarse, int fc,
                           hipStream_t s) {
  if (cg) {
    if (coarse) k_down<T, NU, true, true, FC0><<<tiles(a.F, FC0), BS, 0, s>>>(a);
    else k_down<T, NU, true, false, FC0><<<tiles(a.F, FC0), BS, 0, s>>>(a);
  } else {
    if (coarse) launch_down_fc<T, NU, false, true>(a, fc, s);
    else launch_down_fc<T, NU, false, false>(a, fc, s);
  }
}

template <class T>
static void launch_down(const DownArgs<T>& a, int nu, bool cg, bool coarse,
                        int fc, hipStream_t s) {
  if (nu == 1) launch_down_nu<T, 1>(a, cg, coarse, fc, s);
  else launch_down_nu<T, 2>(a, cg, coarse, fc, s);
}

template <class T, int NU, bool FINE, bool COARSE>
static void launch_up_fc(const UpArgs<T>& a, int fc, hipStream_t s) {
  if (fc == 16)
    k_up<T, NU, FINE, COARSE, 16><<<tiles(a.F, 16), BS, 0, s>>>(a);
  else
    k_up<T, NU, FINE, COARSE, 32><<<tiles(a.F, 32), BS, 0, s>>>(a);
}

template <class T, int NU>
static void launch_up_nu(const UpArgs<T>& a, bool fine, bool coarse, int fc,
                         hipStream_t s) {
  if (fine) {
    if (coarse) k_up<T, NU, true, true, FC0><<<tiles(a.F, FC0), BS, 0, s>>>(a);
    else k_up<T, NU, true, false, FC0><<<tiles(a.F, FC0), BS, 0, s>>>(a);
  } else {
    if (coarse) launch_up_fc<T, NU, false, true>(a, fc, s);
    else launch_up_fc<T, NU, false, false>(a, fc, s);
  }
}

template <class T>
static void launch_up(const UpArgs<T>& a, int nu, bool fine, bool coarse,
                      int fc, hipStream_t s) {
  if (nu == 1) launch_up_nu<T, 1>(a, fine, coarse, fc, s);
  else launch_up_nu<T, 2>(a, fine, coarse, fc, s);
}

// tail workgroup size: opts.tail_threads (256, 512 or 1024; default 1024)
template <class T>
static void launch_tail(f2d_poisson* h, const TailArgs<T>& ta, size_t lds,
                        hipStream_t s) {
  int tb = h->opts.tail_threads;
  if (tb != 256 && tb != 512) tb = 1024;
  k_tail<T><<<1, tb, lds, s>>>(ta);
}

static int nb0(const LevelDev& F) {
  dim3 g = tiles(F, FC0);
  return (int)(g.x * g.y);
}

// One PCG iteration k (k = 0: preconditioning of the initial residual).
// `last`: only the update and its residual (the iteration cap was reached).
template <class T>
static void pcg_iteration(f2d_poisson* h, Pcg<T>& P, int k, T* x,
                          const Ctl* outer, bool inner, bool last,
                          hipStream_t s) {
  const int nu = h->opts.nu, nuc = h->nu_coarse;
  const LevelDev& F = P.lv[0].d;
  const bool has_coarse = P.L > 1;
  const int nb = nb0(F);
  T* r_new = P.r[(k + 1) & 1];
  {
    DownArgs<T> a{};
    a.F = F;
    if (has_coarse) a.C = P.lv[1].d;
    a.r_in = P.r[k & 1];
    a.r_out = r_new;
    a.q = P.q;
    a.p = P.p[(k + 1) & 1];  // p written by iteration k-1
    a.x = x;
    a.zpre = (T*)P.lv[0].z;
    a.rc = has_coarse ? (T*)P.lv[1].r : nullptr;
    a.ctl = P.ctl;
    a.outer = outer;
    a.part_in = h->reg(R_PQ);
    a.part_out = h->reg(R_RR);
    a.nb_in = nb;
    a.first = (k == 0);
    a.iter = k;
    a.inner = inner;
    a.singular = h->singular;
    launch_down<T>(a, nu, true, has_coarse, FC0, s);
  }
  if (last) return;
  const int Lm = P.Lt;  // multi-block levels 0 .. Lt-1
  for (int l = 1; l < Lm; l++) {
    DownArgs<T> a{};
    a.F = P.lv[l].d;
    bool c = l + 1 < P.L;
    if (c) a.C = P.lv[l + 1].d;
    a.r_in = (const T*)P.lv[l].r;
    a.zpre = (T*)P.lv[l].z;
    a.rc = c ? (T*)P.lv[l + 1].r : nullptr;
    a.ctl = P.ctl;
    a.outer = outer;
    launch_down<T>(a, nuc, false, c, P.lv[l].fc, s);
  }
  if (P.Lt < P.L) {
    TailArgs<T> ta = P.tail;
    ta.outer = outer;
    launch_tail<T>(h, ta, P.tail_lds, s);
  }
  for (int l = Lm - 1; l >= 1; l--) {
    UpArgs<T> a{};
    a.F = P.lv[l].d;
    bool c = l + 1 < P.L;
    if (c) a.C = P.lv[l + 1].d;
    a.r = (const T*)P.lv[l].r;
    a.zpre = (const T*)P.lv[l].z;
    a.ec = c ? (const T*)P.lv[l + 1].e : nullptr;
    a.out = (T*)P.lv[l].e;
    a.ctl = P.ctl;
    a.outer = outer;
    launch_up<T>(a, nuc, false, c, P.lv[l].fc, s);
  }
  {
    UpArgs<T> a{};
    a.F = F;
    if (has_coarse) a.C = P.lv[1].d;
    a.r = r_new;
    a.zpre = (const T*)P.lv[0].z;
    a.ec = has_coarse ? (const T*)P.lv[1].e : nullptr;
    a.out = P.z;
    a.ctl = P.ctl;
    a.outer = outer;
    a.part_in = h->reg(R_RR);
    a.part_out = h->reg(R_RZ);
    a.nb_in = nb;
    launch_up<T>(a, nu, true, has_coarse, FC0, s);
  }
  {
    PArgs<T> a{};
    a.L = F;
    a.z = P.z;
    a.p_old = P.p[(k + 1) & 1];
    a.p = P.p[k & 1];
    a.q = P.q;
    a.ctl = P.ctl;
    a.outer = outer;
    a.part_in = h->reg(R_RZ);
    a.part_out = h->reg(R_PQ);
    a.nb_in = nb;
    a.first = (k == 0);
    a.iter = k;
    a.singular = h->singular;
    k_pupdate<T><<<tiles(F, FC0), BS, 0, s>>>(a);
  }
}

template <class T>
static void pcg_loop(f2d_poisson* h, Pcg<T>& P, T* x, int maxit,
                     bool sync_checks, hipStream_t s) {
  int est = h->last_iters > 0 ? h->last_iters : 1;
  for (int k = 0; k <= maxit; k++) {
    pcg_iteration<T>(h, P, k, x, nullptr, false, k == maxit, s);
    if (sync_checks && k + 1 >= est && k < maxit) {
      int done = 0;
      F2D_HIP(hipMemcpyAsync(&done, &P.ctl->done, sizeof(int),
                             hipMemcpyDeviceToHost, s));
      F2D_HIP(hipStreamSynchronize(s));
      if (done) break;
    }
  }
  k_finish<<<1, BS, 0, s>>>(P.ctl, h->reg(R_RR), nb0(P.lv[0].d));
}

template <class T, class TB>
static void solve_plain(f2d_poisson* h, Pcg<T>& P, const TB* b, T* x,
                        int gmode, double rtol, int maxit, bool sync,
                        hipStream_t s) {
  const LevelDev& F = P.lv[0].d;
  InitArgs<T, TB> ia{};
  ia.L = F;
  ia.b = b;
  ia.x = x;
  ia.r = P.r[0];
  ia.ctl = P.ctl;
  ia.part = h->reg(R_INIT);
  ia.rtol = rtol;
  dim3 g = tiles(F, FC0);
  const int nb = nb0(F);
  if (gmode == G_FROM_B)
    k_init<T, TB, G_FROM_B><<<g, BS, 0, s>>>(ia);
  else
    k_init<T, TB, G_FROM_C2V><<<g, BS, 0, s>>>(ia);
  k_init_fin<<<1, BS, 0, s>>>(P.ctl, h->reg(R_INIT), nb, h->singular);
  pcg_loop<T>(h, P, x, maxit, sync, s);
  if (h->singular) {
    k_sum<T><<<g, BS, 0, s>>>(F, x, h->reg(R_MEAN));
    k_mean_fin<<<1, BS, 0, s>>>(h->reg(R_MEAN), nb, h->nwet, h->scal);
    k_sub_mean<T><<<grid1d((int64_t)F.ny * F.nx, BS), BS, 0, s>>>(F, x, h->scal);
  }
}

template <class TB, class TX>
static void solve_refine(f2d_poisson* h, const TB* b, TX* x, int gmode,
                         double rtol, int maxit, bool sync, hipStream_t s) {
  Pcg<float>& P = h->p32;
  const LevelDev& F = P.lv[0].d;
  const int npt = grid1d((int64_t)F.ny * F.nx, BS);
  dim3 g = tiles(F, FC0);
  const int nb = nb0(F);
  if (!h->have_x64) {
    // seed the FP64 iterate from the caller's array (buffer 0 is current)
    k_refine_seed<TX><<<npt, BS, 0, s>>>(F, x, h->x64[0]);
    h->have_x64 = 1;
  }
  k_reset_outer<<<1, 64, 0, s>>>(h->outer, h->nwet, rtol, h->inner_rtol);
  if (h->singular) {
    k_gsum<TB><<<g, BS, 0, s>>>(F, b, gmode, h->reg(R_MEAN));
    k_mean_fin<<<1, BS, 0, s>>>(h->reg(R_MEAN), nb, h->nwet, h->scal + 1);
  }
  const int inner_it = h->opts.refine_inner > 0 ? h->opts.refine_inner : 4;
  const int nref = maxit > 0 ? maxit : 4;
  for (int it = 0; it <= nref; it++) {
    ResArgs<TB> ra{};
    ra.L = F;
    ra.b = b;
    ra.x64a = h->x64[0];
    ra.x64b = h->x64[1];
    ra.e = h->e32;
    ra.r32 = P.r[0];
    ra.inner = P.ctl;
    ra.outer = h->outer;
    ra.part = h->reg(R_REF);
    ra.gmode = gmode;
    ra.singular = h->singular;
    ra.add_e = it > 0;
    ra.gmean = h->scal + 1;
    k_refine_res<TB><<<g, BS, 0, s>>>(ra);
    k_refine_check<<<1, BS, 0, s>>>(h->outer, P.ctl, h->reg(R_REF), nb,
                                    h->singular, it == nref);
    if (it == nref) break;
    // inner FP32 solve of K e = r32 from e = 0 (at most inner_it
    // iterations), predicated off once the outer loop has converged
    for (int k = 0; k <= inner_it; k++)
      pcg_iteration<float>(h, P, k, h->e32, h->outer, true, k == inner_it, s);
    if (sync && it >= 1) {
      int od = 0;
      F2D_HIP(hipMemcpyAsync(&od, &h->outer->odone, sizeof(int),
                             hipMemcpyDeviceToHost, s));
      F2D_HIP(hipStreamSynchronize(s));
      if (od) break;
    }
  }
  const double* shift = nullptr;
  if (h->singular) {
    k_sum64<<<g, BS, 0, s>>>(F, h->x64[0], h->x64[1], h->outer, h->reg(R_MEAN));
    k_mean_fin<<<1, BS, 0, s>>>(h->reg(R_MEAN), nb, h->nwet, h->scal);
    shift = h->scal;
  }
  k_refine_out<TX><<<npt, BS, 0, s>>>(F, h->x64[0], h->x64[1], h->outer, shift,
                                      x);
  k_finish_refine<<<1, 64, 0, s>>>(P.ctl, h->outer);
}

template <class T>
static void fill_x(f2d_poisson* h, T* x, hipStream_t s) {
  if (!h->xper && !h->yper) return;
  int64_t n = h->NYP * h->NXP;
  k_fill_x<T><<<grid1d(n, 256), 256, 0, s>>>(x, (int)h->ny, (int)h->nx,
                                             h->xper, h->yper);
}

template <class TB>
static void do_solve(f2d_poisson* h, const TB* b, void* x, int gmode,
                     double rtol, int maxit, bool sync, hipStream_t s) {
  const int mode = h->opts.mode;
  if (mode == F2D_POISSON_F64) {
    solve_plain<double, TB>(h, h->p64, b, (double*)x, gmode, rtol, maxit, sync, s);
    fill_x<double>(h, (double*)x, s);
  } else if (mode == F2D_POISSON_F32) {
    solve_plain<float, TB>(h, h->p32, b, (float*)x, gmode, rtol, maxit, sync, s);
    fill_x<float>(h, (float*)x, s);
  } else {
    solve_refine<TB, TB>(h, b, (TB*)x, gmode, rtol, maxit, sync, s);
    fill_x<TB>(h, (TB*)x, s);
  }
}

// Launch one fine-level kernel of the PCG iteration `reps` times on the
// handle's current vectors, with a scratch control block and scratch
// outputs, and return the mean launch time (HIP events on `s`).
template <class T>
static double bench_kernel(f2d_poisson* h, Pcg<T>& P, int kind, int reps,
                           hipStream_t s, double* bytes) {
  const int nu = h->opts.nu;
  const LevelDev& F = P.lv[0].d;
  const bool has_coarse = P.L > 1;
  const int nb = nb0(F);
  Ctl c0{};
  c0.nwet = h->nwet;
  c0.bb = 1.0;
  c0.rtol = -1.0;
  c0.rz[0] = c0.rz[1] = 1.0;
  F2D_HIP(hipMemcpy(h->bench_ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice));
  const double pts = (double)F.ny * F.nx;
  const double t = sizeof(T);
  hipEvent_t e0, e1;
  F2D_HIP(hipEventCreate(&e0));
  F2D_HIP(hipEventCreate(&e1));
  double* scratch_part = h->reg(R_MEAN);
  auto launch = [&]() {
    if (kind == 0) {
      DownArgs<T> a{};
      a.F = F;
      if (has_coarse) a.C = P.lv[1].d;
      a.r_in = P.r[0];
      a.r_out = P.r[1];
      a.q = P.q;
      a.p = P.p[0];
      a.x = P.z;  // scratch iterate
      a.zpre = (T*)P.lv[0].z;
      a.rc = has_coarse ? (T*)P.lv[1].r : nullptr;
      a.ctl = h->bench_ctl;
      a.part_in = h->reg(R_PQ);
      a.part_out = scratch_part;
      a.nb_in = nb;
      a.first = 0;
      a.iter = 2;
      a.singular = h->singular;
      launch_down<T>(a, nu, true, has_coarse, FC0, s);
    } else if (kind == 1) {
      UpArgs<T> a{};
      a.F = F;
      if (has_coarse) a.C = P.lv[1].d;
      a.r = P.r[1];
      a.zpre = (const T*)P.lv[0].z;
      a.ec = has_coarse ? (const T*)P.lv[1].e : nullptr;
      a.out = P.q;  // scratch output
      a.ctl = h->bench_ctl;
      a.part_in = h->reg(R_RR);
      a.part_out = scratch_part;
      a.nb_in = nb;
      a.bench = 1;
      launch_up<T>(a, nu, true, has_coarse, FC0, s);
    } else {
      PArgs<T> a{};
      a.L = F;
      a.z = P.z;
      a.p_old = P.p[0];
      a.p = P.p[1];
      a.q = P.q;
      a.ctl = h->bench_ctl;
      a.part_in = h->reg(R_RZ);
      a.part_out = scratch_part;
      a.nb_in = nb;
      a.first = 0;
      a.iter = 2;
      a.singular = h->singular;
      a.bench = 1;
      k_pupdate<T><<<tiles(F, FC0), BS, 0, s>>>(a);
    }
  };
  // algorithmic bytes: each distinct array read once and written once
  if (kind == 0)  // r, q, p, x read; r, x, zpre written; rc (1/4); mask
    *bytes = pts * (7 * t + (has_coarse ? 0.25 * t : 0.0) + 1.0);
  else if (kind == 1)  // r, zpre, e_c (1/4) read; z written; mask
    *bytes = pts * (3 * t + (has_coarse ? 0.25 * t : 0.0) + 1.0);
  else  // z, p_old read; p, q written; mask
    *bytes = pts * (4 * t + 1.0);
  launch();
  F2D_HIP(hipStreamSynchronize(s));
  F2D_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; i++) launch();
  F2D_HIP(hipEventRecord(e1, s));
  F2D_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  F2D_HIP(hipEventElapsedTime(&ms, e0, e1));
  F2D_HIP(hipEventDestroy(e0));
  F2D_HIP(hipEventDestroy(e1));
  return 1e3 * (double)ms / reps;
}

}  // namespace f2d

using namespace f2d;

extern "C" {

int f2d_poisson_bench_kernel(f2d_poisson* h, int kind, int reps, double* us,
                             double* bytes, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(h && us && bytes, "NULL argument");
    F2D_REQUIRE(!h->neu, "no kernel timing on a Neumann (Jacobi-PCG) handle");
    F2D_REQUIRE(kind >= 0 && kind <= 9, "kind must be in 0 .. 9");
    F2D_REQUIRE(reps > 0, "reps must be > 0");
    hipStream_t s = S(stream);
    if (kind >= 3) {
      F2D_REQUIRE(h->fft, "FFT kernels timed on a handle without the FFT solve");
      *us = fft_bench(h->fft, h->opts.mode == F2D_POISSON_F64, kind - 3, reps, s, bytes);
      return;
    }
    F2D_REQUIRE(!h->fft || h->p64.L > 0 || h->p32.L > 0, "no multigrid hierarchy");
    if (h->opts.mode == F2D_POISSON_F64)
      *us = bench_kernel<double>(h, h->p64, kind, reps, s, bytes);
    else
      *us = bench_kernel<float>(h, h->p32, kind, reps, s, bytes);
  });
}

int f2d_poisson_create(const int8_t* unknowns_host, const f2d_grid* g,
                       double maindiag, const f2d_poisson_opts* opts,
                       f2d_poisson** out) {
  f2d_poisson* h = nullptr;
  int rc = guarded([&] {
    F2D_REQUIRE(out, "out is NULL");
    F2D_REQUIRE(g && g->halowidth == 3, "grid: halowidth must be 3");
    F2D_REQUIRE(g->ny > 0 && g->nx > 0, "grid: empty");
    F2D_REQUIRE(g->dx > 0 && g->dy > 0, "grid: dx, dy must be positive");
    F2D_REQUIRE(maindiag >= 0, "maindiag must be >= 0");
    h = new f2d_poisson();
    h->ny = g->ny;
    h->nx = g->nx;
    h->NYP = g->ny + 6;
    h->NXP = g->nx + 6;
    h->xper = g->xperiodic;
    h->yper = g->yperiodic;
    h->maindiag = maindiag;
    h->opts.mode = F2D_POISSON_F64;
    h->opts.nu = 2;
    h->opts.refine_inner = 4;
    h->opts.tail_points = 1024;
    if (opts) {
      h->opts = *opts;
      if (h->opts.nu <= 0) h->opts.nu = 2;
      if (h->opts.refine_inner <= 0) h->opts.refine_inner = 4;
      if (h->opts.tail_points <= 0) h->opts.tail_points = 1024;
    }
    F2D_REQUIRE(h->opts.nu == 1 || h->opts.nu == 2, "opts.nu must be 1 or 2");
    h->nu_coarse = h->opts.nu_coarse > 0 ? h->opts.nu_coarse : h->opts.nu;
    h->nu_tail = h->opts.nu_tail > 0 ? h->opts.nu_tail : h->opts.nu;
    F2D_REQUIRE(h->nu_coarse <= 2 && h->nu_tail <= 2,
                "coarse/tail sweeps must be 1 or 2");
    F2D_REQUIRE(h->opts.mode >= 0 && h->opts.mode <= 2, "opts.mode unknown");
    // unknowns must lie in the interior block [3, ny+3) x [3, nx+3)
    size_t N = (size_t)(h->NYP * h->NXP);
    std::vector<int8_t> m(N, 0);
    int64_t nwet = 0;
    for (int64_t j = 0; j < h->NYP; j++)
      for (int64_t i = 0; i < h->NXP; i++) {
        int8_t v = unknowns_host[j * h->NXP + i] ? 1 : 0;
        bool inside = j >= 3 && j < 3 + h->ny && i >= 3 && i < 3 + h->nx;
        F2D_REQUIRE(!v || inside,
                    "Poisson unknowns outside the interior block (wet halo "
                    "in a closed direction) are not supported");
        m[j * h->NXP + i] = v;
        nwet += v;
      }
    h->nwet = (double)nwet;
    // the Neumann cell operator: the direct solve when every cell is wet
    // (closed directions as even extensions), Jacobi-PCG otherwise
    const bool fits = fft_fits(h->ny, h->nx, h->xper, h->yper,
                               h->opts.mode == F2D_POISSON_F64);
    const bool cells_fft = h->opts.neumann && fits &&
                           fft_applicable_cells(h->ny, h->nx, h->xper, h->yper, nwet);
    if (h->opts.neumann && !(cells_fft && h->opts.solver != F2D_SOLVER_MG)) {
      F2D_REQUIRE(h->opts.solver != F2D_SOLVER_FFT,
                  "FFT solver requested for a Neumann operator with dry cells or sizes "
                  "that are not powers of two");
      h->neu = neu_create(m.data(), h->ny, h->nx, h->xper, h->yper, g->dx, g->dy, maindiag);
      *out = h;
      return;
    }
    h->singular = (maindiag == 0.0 && h->xper && h->yper &&
                   nwet == h->ny * h->nx)
                      ? 1
                      : 0;
    h->m0 = (int8_t*)dalloc(h, N);
    F2D_HIP(hipMemcpy(h->m0, m.data(), N, hipMemcpyHostToDevice));
    LevelDev& b = h->base;
    b.ny = (int)h->ny;
    b.nx = (int)h->nx;
    b.stride = (int)h->NXP;
    b.origin = 3 * h->NXP + 3;
    b.xper = h->xper;
    b.yper = h->yper;
    b.cx = g->dy / g->dx;
    b.cy = g->dx / g->dy;
    b.d = 2.0 * (b.cx + b.cy) + maindiag;
    b.inv_d = 1.0 / b.d;
    b.m = h->m0;
    // partial-sum regions: REGION_SLOTS values per level-0 tile
    int64_t ntile = ((h->ny + FC0 - 1) / FC0) * ((h->nx + FC0 - 1) / FC0);
    h->region = REGION_SLOTS * (ntile + 64);
    h->part = (double*)dalloc(h, sizeof(double) * R_COUNT * h->region);
    h->scal = (double*)dalloc(h, sizeof(double) * 8);
    h->outer = (Ctl*)dalloc(h, sizeof(Ctl));
    h->bench_ctl = (Ctl*)dalloc(h, sizeof(Ctl));
    if (h->opts.mode == F2D_POISSON_F64) {
      build_pcg<double>(h, h->p64);
      if (h->p64.Lt < h->p64.L) set_tail_attr<double>(h->p64.tail_lds);
    } else {
      build_pcg<float>(h, h->p32);
      if (h->p32.Lt < h->p32.L) set_tail_attr<float>(h->p32.tail_lds);
      if (h->opts.mode == F2D_POISSON_F32_REFINE) {
        h->x64[0] = (double*)dalloc(h, N * sizeof(double));
        h->x64[1] = (double*)dalloc(h, N * sizeof(double));
        h->e32 = (float*)dalloc(h, N * sizeof(float));
      }
    }
    const bool fft_ok =
        h->opts.neumann ? cells_fft
                        : fits && fft_applicable(h->ny, h->nx, h->xper, h->yper, nwet);
    F2D_REQUIRE(h->opts.solver != F2D_SOLVER_FFT || fft_ok,
                "FFT solver requested but the operator is not fully wet with "
                "power-of-two sizes (<= 4096, <= 2048 in a closed direction; "
                "FP64: transformed sizes <= 2048)");
    F2D_REQUIRE(h->opts.solver >= 0 && h->opts.solver <= F2D_SOLVER_CAP &&
                    h->opts.solver != F2D_SOLVER_NEUMANN_CG,
                "opts.solver unknown");
    if (fft_ok && h->opts.solver != F2D_SOLVER_MG && h->opts.solver != F2D_SOLVER_CAP)
      h->fft = fft_create(h->ny, h->nx, g->dx, g->dy, maindiag,
                          h->opts.mode == F2D_POISSON_F64,
                          h->opts.mode != F2D_POISSON_F64, h->xper, h->yper,
                          h->opts.neumann != 0);
    // a masked vertex operator inside a rectangle the spectral solve
    // handles: the direct capacitance solve (auto: with a boundary of at
    // most CAP_AUTO_MAX points, and multigrid if its set-up fails)
    if (!h->fft && !h->opts.neumann && h->opts.solver != F2D_SOLVER_MG &&
        h->opts.mode != F2D_POISSON_F32) {
      const bool f64m = h->opts.mode == F2D_POISSON_F64;
      const int64_t nb = cap_boundary(m.data(), h->ny, h->nx, h->xper, h->yper, maindiag, f64m);
      if (h->opts.solver == F2D_SOLVER_CAP) {
        F2D_REQUIRE(nb > 0 && nb <= CAP_MAX,
                    "capacitance solver requested but the operator is not a masked "
                    "vertex operator inside a rectangle the spectral solve handles "
                    "(power-of-two sizes, a closed direction or maindiag > 0, at most " +
                        std::to_string(CAP_MAX) + " boundary points)");
        h->cap = cap_create(m.data(), h->ny, h->nx, g->dx, g->dy, maindiag, h->xper,
                            h->yper, f64m);
      } else if (nb > 0 && nb <= CAP_AUTO_MAX) {
        try {
          h->cap = cap_create(m.data(), h->ny, h->nx, g->dx, g->dy, maindiag, h->xper,
                              h->yper, f64m);
        } catch (const Error&) {
          h->cap = nullptr;  // multigrid solves this operator
          (void)hipGetLastError();
        }
      }
    }
    F2D_HIP(hipDeviceSynchronize());
    *out = h;
  });
  if (rc != 0 && h) {
    if (h->fft) fft_destroy(h->fft);
    if (h->cap) cap_destroy(h->cap);
    if (h->neu) neu_destroy(h->neu);
    for (void* p : h->allocs) (void)hipFree(p);
    delete h;
  }
  return rc;
}

int f2d_poisson_destroy(f2d_poisson* h) {
  return guarded([&] {
    if (!h) return;
    F2D_HIP(hipDeviceSynchronize());
    if (h->fft) fft_destroy(h->fft);
    if (h->cap) cap_destroy(h->cap);
    if (h->neu) neu_destroy(h->neu);
    for (void* p : h->allocs) F2D_HIP(hipFree(p));
    delete h;
  });
}

static int solve_common(f2d_poisson* h, const void* b, void* x, int gmode,
                        double rtol, int maxit, int* iters, double* relres,
                        void* stream) {
  return guarded([&] {
    F2D_REQUIRE(h, "handle is NULL");
    F2D_REQUIRE(b && x, "b/x is NULL");
    F2D_REQUIRE(maxit >= 0, "maxit must be >= 0");
    F2D_REQUIRE(rtol >= 0, "rtol must be >= 0");
    hipStream_t s = S(stream);
    bool sync = iters != nullptr || relres != nullptr;
    const int mode = h->opts.mode;
    F2D_REQUIRE(!h->opts.neumann || gmode == G_FROM_B,
                "the Neumann operator takes b on the cells");
    if (h->neu) {
      int it = 0;
      double rr = 0.0;
      neu_solve(h->neu, mode == F2D_POISSON_F64 ? F2D_F64 : F2D_F32, b, x, rtol, maxit, sync,
                &it, &rr, s);
      if (sync) {
        if (iters) *iters = it;
        if (relres) *relres = rr;
        h->last_iters = it;
      }
      return;
    }
    if (h->cap) {
      Ctl* c = mode == F2D_POISSON_F64 ? h->p64.ctl : h->p32.ctl;
      if (mode == F2D_POISSON_F64) {
        cap_solve_f64(h->cap, (const double*)b, gmode, (double*)x, sync, c->stats, s);
        fill_x<double>(h, (double*)x, s);
      } else {
        const int nref = h->refine_next >= 0
                             ? h->refine_next
                             : (h->opts.fft_refine > 0 ? h->opts.fft_refine : 1);
        cap_solve_f32(h->cap, (const float*)b, gmode, (float*)x, nref,
                      h->opts.fft_warm != 0, sync, c->stats, s);
        fill_x<float>(h, (float*)x, s);
      }
      if (!sync) h->fft_unmeasured++;
    } else if (h->fft) {
      Ctl* c = mode == F2D_POISSON_F64 ? h->p64.ctl : h->p32.ctl;
      if (mode == F2D_POISSON_F64) {
        fft_solve_f64(h->fft, (const double*)b, gmode, (double*)x, rtol, sync,
                      c->stats, s);
      } else {
        const int nref = mode == F2D_POISSON_F32
                             ? 0
                             : (h->refine_next >= 0
                                    ? h->refine_next
                                    : (h->opts.fft_refine > 0 ? h->opts.fft_refine : 1));
        fft_solve_f32_refine(h->fft, (const float*)b, gmode, (float*)x, nref,
                             h->opts.fft_warm != 0, rtol,
                             sync, c->stats, s);
      }
      if (!sync) h->fft_unmeasured++;
    } else if (mode == F2D_POISSON_F64) {
      do_solve<double>(h, (const double*)b, x, gmode, rtol, maxit, sync, s);
    } else {
      do_solve<float>(h, (const float*)b, x, gmode, rtol, maxit, sync, s);
    }
    if (sync) {
      Ctl* c = mode == F2D_POISSON_F64 ? h->p64.ctl : h->p32.ctl;
      double st[8];
      F2D_HIP(hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, s));
      F2D_HIP(hipStreamSynchronize(s));
      if (iters) *iters = (int)st[0];
      if (relres) *relres = st[1];
      if (mode != F2D_POISSON_F32_REFINE) h->last_iters = (int)st[0];
    }
  });
}

int f2d_poisson_solve(f2d_poisson* h, const void* b, void* x, double rtol,
                      int maxit, int* iters, double* relres, void* stream) {
  return solve_common(h, b, x, G_FROM_B, rtol, maxit, iters, relres, stream);
}

int f2d_poisson_solve_c2v(f2d_poisson* h, const void* omega, void* x,
                          double rtol, int maxit, int* iters, double* relres,
                          void* stream) {
  return solve_common(h, omega, x, G_FROM_C2V, rtol, maxit, iters, relres,
                      stream);
}

int f2d_poisson_set_refine(f2d_poisson* h, int nref) {
  return guarded([&] {
    F2D_REQUIRE(h, "handle is NULL");
    F2D_REQUIRE(nref >= -1 && nref <= 16, "refinements must be in -1 .. 16");
    h->refine_next = nref;
  });
}

int f2d_poisson_method(f2d_poisson* h) {
  if (!h) return -1;
  if (h->neu) return F2D_SOLVER_NEUMANN_CG;
  if (h->cap) return F2D_SOLVER_CAP;
  return h->fft ? F2D_SOLVER_FFT : F2D_SOLVER_MG;
}

int f2d_poisson_reset_stats(f2d_poisson* h, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(h, "handle is NULL");
    h->fft_unmeasured = 0;
    if (h->neu) {
      neu_reset_stats(h->neu, S(stream));
      return;
    }
    Ctl* c = h->opts.mode == F2D_POISSON_F64 ? h->p64.ctl : h->p32.ctl;
    k_reset_stats<<<1, 64, 0, S(stream)>>>(c);
  });
}

int f2d_poisson_read_stats(f2d_poisson* h, double* stats_host, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(h && stats_host, "NULL argument");
    const double* src = h->neu ? neu_stats(h->neu)
                               : (h->opts.mode == F2D_POISSON_F64 ? h->p64.ctl : h->p32.ctl)
                                     ->stats;
    F2D_HIP(hipMemcpyAsync(stats_host, src, 8 * sizeof(double),
                           hipMemcpyDeviceToHost, S(stream)));
    F2D_HIP(hipStreamSynchronize(S(stream)));
    // direct solves made without statistics count as converged solves
    stats_host[6] += (double)h->fft_unmeasured;
  });
}

}  // extern "C"

struct f2d_poisson_batch {
  std::vector<f2d_poisson*> h;
  std::vector<void*> x;
  f2d::CapBatch* cap = nullptr;
  int nref = 1;
  bool warm = true;
};

extern "C" {

int f2d_poisson_batch_create(f2d_poisson* const* h, int32_t n, const void* const* b,
                             void* const* x, f2d_poisson_batch** out) {
  return guarded([&] {
    F2D_REQUIRE(h && b && x && out, "NULL argument");
    F2D_REQUIRE(n >= 1, "empty batch");
    auto* bt = new f2d_poisson_batch();
    try {
      std::vector<f2d::CapSolver*> c(n);
      for (int i = 0; i < n; i++) {
        F2D_REQUIRE(h[i] && h[i]->cap && h[i]->opts.mode != F2D_POISSON_F64,
                    "batched solves take FP32 capacitance solvers");
        F2D_REQUIRE(h[i]->opts.fft_refine == h[0]->opts.fft_refine &&
                        (h[i]->opts.fft_warm != 0) == (h[0]->opts.fft_warm != 0),
                    "batch members must share the refinement options");
        bt->h.push_back(h[i]);
        bt->x.push_back(x[i]);
        c[i] = h[i]->cap;
      }
      bt->nref = h[0]->opts.fft_refine > 0 ? h[0]->opts.fft_refine : 1;
      bt->warm = h[0]->opts.fft_warm != 0;
      bt->cap = f2d::cap_batch_create(c.data(), n, reinterpret_cast<const float* const*>(b),
                                      reinterpret_cast<float* const*>(x));
    } catch (...) {
      delete bt;
      throw;
    }
    *out = bt;
  });
}

int f2d_poisson_batch_solve(f2d_poisson_batch* bt, int nref, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(bt, "batch is NULL");
    F2D_REQUIRE(nref >= -1 && nref <= 16, "refinements must be in -1 .. 16");
    hipStream_t s = S(stream);
    f2d::cap_batch_solve(bt->cap, nref >= 0 ? nref : bt->nref, bt->warm, s);
    for (size_t i = 0; i < bt->h.size(); i++) {
      f2d::fill_x<float>(bt->h[i], (float*)bt->x[i], s);
      bt->h[i]->fft_unmeasured++;
    }
  });
}

int f2d_poisson_batch_destroy(f2d_poisson_batch* bt) {
  return guarded([&] {
    if (!bt) return;
    f2d::cap_batch_destroy(bt->cap);
    delete bt;
  });
}

const double* f2d_poisson_stats(f2d_poisson* h) {
  if (!h) return nullptr;
  if (h->neu) return neu_stats(h->neu);
  Ctl* c = h->opts.mode == F2D_POISSON_F64 ? h->p64.ctl : h->p32.ctl;
  return c->stats;
}

}  // extern "C"
