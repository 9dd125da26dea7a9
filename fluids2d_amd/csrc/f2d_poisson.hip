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
// Kernels per CG iteration (level 0 = the fine grid):
//   k_down<CG>   r -= alpha q, x += alpha p, ||r||^2 + convergence test,
//                then pre-smooth / residual / restrict  (one LDS tile pass)
//   k_down       the same smoothing/restriction on levels 1 .. Lt-1
//   k_tail       one workgroup runs the levels Lt .. L-1 entirely in LDS
//   k_up         prolongate + post-smooth on levels Lt-1 .. 1
//   k_up<FINE>   level 0 post-smoothing, (r, z), sum z -> beta
//   k_pupdate    p = z + beta p, q = K p, (p, q) -> alpha
// Scalars live in a device control block; a converged solve turns the rest
// of the enqueued kernels into early returns.
#include <cmath>
#include <cstring>
#include <vector>

#include "f2d_common.hpp"

namespace f2d {

struct LevelDev {
  int ny, nx, stride, xper, yper;
  int64_t origin;
  double cx, cy, d, inv_d;
  const int8_t* m;
};

struct Ctl {
  double bb;      // ||g||^2
  double rr;      // ||r||^2 (latest)
  double rz;      // (r, z) of the latest V-cycle
  double pq;
  double alpha, beta;
  double mean_r;  // mean of the initial residual (singular operator)
  double mean_z;  // mean of z (singular operator)
  double relres;
  double rtol;
  double nwet;
  double pad0;
  int done;
  int iters;
  int x_valid;  // inner solve: the iterate has been written
  int pad1;
  // outer (refinement) control
  double obb, orr, orelres;
  int odone, orefine;
  int xcur;        // which FP64 iterate buffer is current
  int inner_total; // inner iterations summed over the refinements
  double stats[4];  // iterations, relres, converged, refinements
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
  double* partials;
  unsigned* counter;
  int singular;
  double rtol;
};

template <class T, class TB, int GM>
__device__ __forceinline__ double g_at(const InitArgs<T, TB>& a, int64_t k) {
  if (GM == G_FROM_B) return -(double)a.b[k];
  const int64_t P = a.L.stride;
  // centerstovertices, operators.py:131-132
  TB s = TB(0.25) * (((a.b[k - P - 1] + a.b[k - 1]) + a.b[k - P]) + a.b[k]);
  return -(double)s;
}

template <class T, class TB, int GM>
__global__ __launch_bounds__(BS) void k_init(InitArgs<T, TB> a) {
  __shared__ double sh[BS];
  __shared__ int flag;
  const LevelDev& L = a.L;
  int64_t p = (int64_t)blockIdx.x * BS + threadIdx.x;
  double gg = 0.0, rs = 0.0;
  if (p < (int64_t)L.ny * L.nx) {
    int y = (int)(p / L.nx), x = (int)(p - (int64_t)y * L.nx);
    int64_t k = li(L, y, x);
    T rv = T(0);
    if (L.m[k]) {
      double g = g_at<T, TB, GM>(a, k);
      T gv = T(g);
      rv = gv;
      if (a.x) {
        T xc = finite_or_zero(a.x[k]);
        a.x[k] = xc;
        T nb[4];
        const int dy[4] = {0, 0, -1, 1}, dx[4] = {-1, 1, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; q++) {
          int yy = y + dy[q], xx = x + dx[q];
          nb[q] = T(0);
          if (lw(L, yy, xx)) {
            int64_t kk = li(L, yy, xx);
            if (L.m[kk]) nb[q] = finite_or_zero(a.x[kk]);
          }
        }
        T Kx = T(L.d) * xc - T(L.cx) * (nb[0] + nb[1]) - T(L.cy) * (nb[2] + nb[3]);
        rv = gv - Kx;
      }
      gg = (double)gv * (double)gv;
      rs = (double)rv;
    }
    a.r[k] = rv;
  }
  double bg = block_sum<BS>(gg, sh);
  double br = block_sum<BS>(rs, sh);
  if (threadIdx.x == 0) {
    store_partial(a.partials + blockIdx.x, bg);
    store_partial(a.partials + gridDim.x + blockIdx.x, br);
  }
  if (!last_block_arrive<BS>(a.counter, &flag)) return;
  double tg = reduce_partials<BS>(a.partials, gridDim.x, sh, false);
  double tr = reduce_partials<BS>(a.partials + gridDim.x, gridDim.x, sh, false);
  if (threadIdx.x == 0) {
    Ctl* c = a.ctl;
    c->bb = tg;
    c->mean_r = a.singular ? tr / c->nwet : 0.0;
    c->rtol = a.rtol;
    c->done = 0;
    c->iters = 0;
    c->x_valid = 0;
    c->alpha = 0.0;
    c->beta = 0.0;
    c->mean_z = 0.0;
    c->rz = 0.0;
    c->relres = 1.0;
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
  double* partials;
  unsigned* counter;
  int first;    // CG: iteration 0 (no update; singular projection)
  int iter;
  int inner;    // inner solve of a refinement: first update writes x
};

template <class T, int NU, bool CG, bool COARSE>
__global__ __launch_bounds__(BS) void k_down(DownArgs<T> a) {
  constexpr int CT = 16, FC = 2 * CT, R = 2 * NU + 1, W = FC + 2 * R;
  __shared__ T sr[W * W];
  __shared__ T sz[W * W];
  __shared__ int8_t sm[W * W];
  __shared__ double sh[BS];
  __shared__ int flag;
  if (stop(a.ctl, a.outer)) return;
  const LevelDev& F = a.F;
  const int t = threadIdx.x;
  const int cy0 = blockIdx.y * CT, cx0 = blockIdx.x * CT;
  const int fy0 = 2 * cy0, fx0 = 2 * cx0;
  T alpha = T(0), shift = T(0);
  bool zero_x = false;
  if (CG) {
    if (a.first) {
      shift = T(a.ctl->mean_r);
      zero_x = a.ctl->bb == 0.0;
    } else {
      alpha = T(a.ctl->alpha);
    }
  }
  for (int k = t; k < W * W; k += BS) {
    int ly = k / W, lx = k - ly * W;
    int y = fy0 - R + ly, x = fx0 - R + lx;
    T rv = T(0);
    int8_t mv = 0;
    if (lw(F, y, x)) {
      int64_t g = li(F, y, x);
      mv = F.m[g];
      if (mv) {
        rv = a.r_in[g];
        if (CG) rv = a.first ? rv - shift : rv - alpha * a.q[g];
      }
    }
    sr[k] = rv;
    sm[k] = mv;
    sz[k] = T(0);
  }
  __syncthreads();

  double rr = 0.0;
  if (CG) {
    for (int e = t; e < FC * FC; e += BS) {
      int cy = e / FC, cx = e - cy * FC;
      int y = fy0 + cy, x = fx0 + cx;
      if (y >= F.ny || x >= F.nx) continue;
      int64_t g = li(F, y, x);
      int k = (cy + R) * W + (cx + R);
      T rv = sr[k];
      a.r_out[g] = rv;
      if (sm[k]) {
        if (zero_x) {
          a.x[g] = T(0);
        } else if (!a.first) {
          T pv = a.p[g];
          a.x[g] = (a.inner && a.iter == 1) ? alpha * pv : a.x[g] + alpha * pv;
        }
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
  if (COARSE) {
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
  const int nb = gridDim.x * gridDim.y;
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (t == 0) store_partial(a.partials + bid, brr);
  if (!last_block_arrive<BS>(a.counter, &flag)) return;
  double tot = reduce_partials<BS>(a.partials, nb, sh, false);
  if (t == 0) {
    Ctl* c = a.ctl;
    c->rr = tot;
    c->iters = a.iter;
    if (!a.first && a.inner) c->x_valid = 1;
    double rel = c->bb > 0.0 ? sqrt(tot / c->bb) : 0.0;
    c->relres = rel;
    if (c->bb == 0.0 || rel <= c->rtol || !(rel == rel)) c->done = 1;
  }
}

// ---------------------------------------------------------------------------
// up pass: x = zpre + P e_coarse, post-smooth; on level 0 also (r,z), sum z
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
  double* partials;
  unsigned* counter;
  int first;
  int singular;
};

template <class T, int NU, bool FINE, bool COARSE>
__global__ __launch_bounds__(BS) void k_up(UpArgs<T> a) {
  constexpr int FC = 32, R = 2 * NU, W = FC + 2 * R;
  constexpr int CW = FC / 2 + 2 * NU + 2;
  __shared__ T sx[W * W];
  __shared__ T sr[W * W];
  __shared__ int8_t sm[W * W];
  __shared__ T sc[COARSE ? CW * CW : 1];
  __shared__ double sh[BS];
  __shared__ int flag;
  if (stop(a.ctl, a.outer)) return;
  const LevelDev& F = a.F;
  const int t = threadIdx.x;
  const int fy0 = blockIdx.y * FC, fx0 = blockIdx.x * FC;
  for (int k = t; k < W * W; k += BS) {
    int ly = k / W, lx = k - ly * W;
    int y = fy0 - R + ly, x = fx0 - R + lx;
    T rv = T(0), zv = T(0);
    int8_t mv = 0;
    if (lw(F, y, x)) {
      int64_t g = li(F, y, x);
      mv = F.m[g];
      if (mv) {
        rv = a.r[g];
        zv = a.zpre[g];
      }
    }
    sr[k] = rv;
    sx[k] = zv;
    sm[k] = mv;
  }
  const int cyb = fy0 / 2 - NU - 1, cxb = fx0 / 2 - NU - 1;
  if (COARSE) {
    const LevelDev& C = a.C;
    for (int k = t; k < CW * CW; k += BS) {
      int ly = k / CW, lx = k - ly * CW;
      int y = cyb + ly, x = cxb + lx;
      T v = T(0);
      if (lw(C, y, x)) v = a.ec[li(C, y, x)];
      sc[k] = v;
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
  double rz = 0.0, zs = 0.0;
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
    }
  }
  if (!FINE) return;
  double brz = block_sum<BS>(rz, sh);
  double bzs = block_sum<BS>(zs, sh);
  const int nb = gridDim.x * gridDim.y;
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (t == 0) {
    store_partial(a.partials + bid, brz);
    store_partial(a.partials + nb + bid, bzs);
  }
  if (!last_block_arrive<BS>(a.counter, &flag)) return;
  double trz = reduce_partials<BS>(a.partials, nb, sh, false);
  double tzs = reduce_partials<BS>(a.partials + nb, nb, sh, false);
  if (t == 0) {
    Ctl* c = a.ctl;
    double mz = a.singular ? tzs / c->nwet : 0.0;
    // (r, z - mz) = (r, z) - mz * sum(r); sum(r) vanishes for the singular
    // operator up to rounding, so (r, z) is kept.
    c->beta = a.first ? 0.0 : trz / c->rz;
    c->rz = trz;
    c->mean_z = mz;
    if (!(trz > 0.0)) c->done = 1;  // breakdown / exact solution
  }
}

// ---------------------------------------------------------------------------
// p = (z - mean z) + beta p_old ;  q = K p ;  alpha = (r,z) / (p,q)
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
  double* partials;
  unsigned* counter;
  int first;
};

template <class T>
__device__ __forceinline__ T pnew_at(const PArgs<T>& a, int y, int x, T beta,
                                     T mz) {
  if (!lw(a.L, y, x)) return T(0);
  int64_t k = li(a.L, y, x);
  if (!a.L.m[k]) return T(0);
  T v = a.z[k] - mz;
  if (!a.first) v = v + beta * a.p_old[k];
  return v;
}

template <class T>
__global__ __launch_bounds__(BS) void k_pupdate(PArgs<T> a) {
  __shared__ double sh[BS];
  __shared__ int flag;
  if (stop(a.ctl, a.outer)) return;
  const LevelDev& L = a.L;
  const T beta = T(a.ctl->beta), mz = T(a.ctl->mean_z);
  int64_t p = (int64_t)blockIdx.x * BS + threadIdx.x;
  double pq = 0.0;
  if (p < (int64_t)L.ny * L.nx) {
    int y = (int)(p / L.nx), x = (int)(p - (int64_t)y * L.nx);
    int64_t k = li(L, y, x);
    T pc = T(0), qv = T(0);
    if (L.m[k]) {
      pc = pnew_at(a, y, x, beta, mz);
      T w = pnew_at(a, y, x - 1, beta, mz), e = pnew_at(a, y, x + 1, beta, mz);
      T s = pnew_at(a, y - 1, x, beta, mz), n = pnew_at(a, y + 1, x, beta, mz);
      qv = T(L.d) * pc - T(L.cx) * (w + e) - T(L.cy) * (s + n);
      pq = (double)pc * (double)qv;
    }
    a.p[k] = pc;
    a.q[k] = qv;
  }
  double b = block_sum<BS>(pq, sh);
  if (threadIdx.x == 0) store_partial(a.partials + blockIdx.x, b);
  if (!last_block_arrive<BS>(a.counter, &flag)) return;
  double tot = reduce_partials<BS>(a.partials, gridDim.x, sh, false);
  if (threadIdx.x == 0) {
    Ctl* c = a.ctl;
    c->pq = tot;
    if (tot > 0.0)
      c->alpha = c->rz / tot;
    else
      c->done = 1;
  }
}

// ---------------------------------------------------------------------------
// tail: levels Lt .. L-1 of the V-cycle in one workgroup, all in LDS
// ---------------------------------------------------------------------------
constexpr int MAXL = 16;
constexpr int TAIL_BS = 1024;

template <class T>
struct TailArgs {
  LevelDev lv[MAXL];
  int off_r[MAXL], off_x[MAXL], off_m[MAXL];
  int nl;
  int coarse_sweeps;
  const T* r_in;
  T* e_out;
  Ctl* ctl;
  const Ctl* outer;
};

template <class T>
__device__ __forceinline__ T lds_at(const T* a, const int8_t* m,
                                    const LevelDev& L, int y, int x) {
  if (!lw(L, y, x)) return T(0);
  int k = y * L.nx + x;
  return m[k] ? a[k] : T(0);
}

template <class T>
__device__ void tail_halfsweep(T* xa, const T* ra, const int8_t* m,
                               const LevelDev& L, int col, bool fromzero) {
  const T cx = T(L.cx), cy = T(L.cy), id = T(L.inv_d);
  const int n = L.ny * L.nx;
  for (int k = threadIdx.x; k < n; k += TAIL_BS) {
    int y = k / L.nx, x = k - y * L.nx;
    if (((y + x) & 1) != col || !m[k]) continue;
    if (fromzero) {
      xa[k] = ra[k] * id;
    } else {
      T w = lds_at(xa, m, L, y, x - 1), e = lds_at(xa, m, L, y, x + 1);
      T s = lds_at(xa, m, L, y - 1, x), nn = lds_at(xa, m, L, y + 1, x);
      xa[k] = (ra[k] + cx * (w + e) + cy * (s + nn)) * id;
    }
  }
  __syncthreads();
}

template <class T>
__device__ __forceinline__ T tail_res(const T* xa, const T* ra,
                                      const int8_t* m, const LevelDev& L,
                                      int y, int x) {
  if (!lw(L, y, x)) return T(0);
  int k = y * L.nx + x;
  if (!m[k]) return T(0);
  T w = lds_at(xa, m, L, y, x - 1), e = lds_at(xa, m, L, y, x + 1);
  T s = lds_at(xa, m, L, y - 1, x), nn = lds_at(xa, m, L, y + 1, x);
  return ra[k] - (T(L.d) * xa[k] - T(L.cx) * (w + e) - T(L.cy) * (s + nn));
}

template <class T, int NU>
__global__ __launch_bounds__(TAIL_BS) void k_tail(TailArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (stop(a.ctl, a.outer)) return;
  const int t = threadIdx.x;
  auto R_ = [&](int l) { return reinterpret_cast<T*>(smem + a.off_r[l]); };
  auto X_ = [&](int l) { return reinterpret_cast<T*>(smem + a.off_x[l]); };
  auto M_ = [&](int l) { return reinterpret_cast<int8_t*>(smem + a.off_m[l]); };
  for (int l = 0; l < a.nl; l++) {
    const LevelDev& L = a.lv[l];
    int n = L.ny * L.nx;
    int8_t* m = M_(l);
    T* r = R_(l);
    T* x = X_(l);
    for (int k = t; k < n; k += TAIL_BS) {
      int y = k / L.nx, xx = k - y * L.nx;
      int64_t g = li(L, y, xx);
      m[k] = L.m[g];
      x[k] = T(0);
      if (l == 0) r[k] = a.r_in[g];
    }
  }
  __syncthreads();
  // down
  for (int l = 0; l + 1 < a.nl; l++) {
    const LevelDev& L = a.lv[l];
    const LevelDev& C = a.lv[l + 1];
    T* x = X_(l);
    T* r = R_(l);
    int8_t* m = M_(l);
    for (int h = 1; h <= 2 * NU; h++)
      tail_halfsweep(x, r, m, L, (h - 1) & 1, h == 1);
    int nc = C.ny * C.nx;
    T* rc = R_(l + 1);
    int8_t* mc = M_(l + 1);
    for (int k = t; k < nc; k += TAIL_BS) {
      int J = k / C.nx, I = k - J * C.nx;
      T v = T(0);
      if (mc[k]) {
        int y = 2 * J, xx = 2 * I;
        v = tail_res(x, r, m, L, y, xx) +
            T(0.5) * ((tail_res(x, r, m, L, y, xx - 1) +
                       tail_res(x, r, m, L, y, xx + 1)) +
                      (tail_res(x, r, m, L, y - 1, xx) +
                       tail_res(x, r, m, L, y + 1, xx))) +
            T(0.25) * ((tail_res(x, r, m, L, y - 1, xx - 1) +
                        tail_res(x, r, m, L, y - 1, xx + 1)) +
                       (tail_res(x, r, m, L, y + 1, xx - 1) +
                        tail_res(x, r, m, L, y + 1, xx + 1)));
        v = T(0.25) * v;
      }
      rc[k] = v;
    }
    __syncthreads();
  }
  // coarsest: palindromic symmetric Gauss-Seidel (red, black)^S (black, red)^S
  {
    int l = a.nl - 1;
    const LevelDev& L = a.lv[l];
    T* x = X_(l);
    T* r = R_(l);
    int8_t* m = M_(l);
    int S = a.coarse_sweeps;
    for (int s = 0; s < S; s++) {
      tail_halfsweep(x, r, m, L, 0, s == 0);
      tail_halfsweep(x, r, m, L, 1, false);
    }
    for (int s = 0; s < S; s++) {
      tail_halfsweep(x, r, m, L, 1, false);
      tail_halfsweep(x, r, m, L, 0, false);
    }
  }
  // up
  for (int l = a.nl - 2; l >= 0; l--) {
    const LevelDev& L = a.lv[l];
    const LevelDev& C = a.lv[l + 1];
    T* x = X_(l);
    T* r = R_(l);
    int8_t* m = M_(l);
    const T* xc = X_(l + 1);
    const int8_t* mc = M_(l + 1);
    int n = L.ny * L.nx;
    for (int k = t; k < n; k += TAIL_BS) {
      if (!m[k]) continue;
      int y = k / L.nx, xx = k - y * L.nx;
      int J = y >> 1, I = xx >> 1;
      T v;
      if (!(y & 1) && !(xx & 1))
        v = lds_at(xc, mc, C, J, I);
      else if (!(y & 1))
        v = T(0.5) * (lds_at(xc, mc, C, J, I) + lds_at(xc, mc, C, J, I + 1));
      else if (!(xx & 1))
        v = T(0.5) * (lds_at(xc, mc, C, J, I) + lds_at(xc, mc, C, J + 1, I));
      else
        v = T(0.25) * ((lds_at(xc, mc, C, J, I) + lds_at(xc, mc, C, J, I + 1)) +
                       (lds_at(xc, mc, C, J + 1, I) +
                        lds_at(xc, mc, C, J + 1, I + 1)));
      x[k] = x[k] + v;
    }
    __syncthreads();
    for (int h = 1; h <= 2 * NU; h++) tail_halfsweep(x, r, m, L, h & 1, false);
  }
  {
    const LevelDev& L = a.lv[0];
    const T* x = X_(0);
    int n = L.ny * L.nx;
    for (int k = t; k < n; k += TAIL_BS) {
      int y = k / L.nx, xx = k - y * L.nx;
      a.e_out[li(L, y, xx)] = x[k];
    }
  }
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

// mean of g over the unknowns, in FP64 (the singular operator's right-hand
// side is projected before any FP32 rounding)
template <class TB>
__global__ __launch_bounds__(BS) void k_gmean(LevelDev L, const TB* b,
                                              int gmode, double* partials,
                                              unsigned* counter, double nwet,
                                              double* out) {
  __shared__ double sh[BS];
  __shared__ int flag;
  int64_t p = (int64_t)blockIdx.x * BS + threadIdx.x;
  double s = 0.0;
  if (p < (int64_t)L.ny * L.nx) {
    int y = (int)(p / L.nx), xx = (int)(p - (int64_t)y * L.nx);
    int64_t k = li(L, y, xx);
    if (L.m[k]) s = g64_at(b, k, L.stride, gmode);
  }
  double bs = block_sum<BS>(s, sh);
  if (threadIdx.x == 0) store_partial(partials + blockIdx.x, bs);
  if (!last_block_arrive<BS>(counter, &flag)) return;
  double tot = reduce_partials<BS>(partials, gridDim.x, sh, false);
  if (threadIdx.x == 0) *out = tot / nwet;
}

template <class TB>
struct ResArgs {
  LevelDev L;
  const TB* b;           // b or omega (padded)
  double* x64a;          // FP64 iterate buffers; outer->xcur selects the
  double* x64b;          // current one, the kernel writes the other
  const float* e;        // inner FP32 correction
  float* r32;            // inner right-hand side (= its initial residual)
  Ctl* inner;
  Ctl* outer;
  double* partials;
  unsigned* counter;
  int gmode, singular, add_e, last;
  double rtol, inner_rtol;
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

template <class TB>
__global__ __launch_bounds__(BS) void k_refine_res(ResArgs<TB> a) {
  __shared__ double sh[BS];
  __shared__ int flag;
  if (a.outer->odone) return;
  const LevelDev& L = a.L;
  const bool addv = a.add_e && a.inner->x_valid;
  const int cur = a.outer->xcur;
  const double* xin = cur ? a.x64b : a.x64a;
  double* xout = cur ? a.x64a : a.x64b;
  int64_t p = (int64_t)blockIdx.x * BS + threadIdx.x;
  double gg = 0.0, rr = 0.0, rs = 0.0, gs = 0.0;
  if (p < (int64_t)L.ny * L.nx) {
    int y = (int)(p / L.nx), x = (int)(p - (int64_t)y * L.nx);
    int64_t k = li(L, y, x);
    float r32 = 0.f;
    if (L.m[k]) {
      double g = g64_at(a.b, k, L.stride, a.gmode);
      if (a.singular) g -= *a.gmean;
      double xc = res_x(a, xin, y, x, addv);
      double w = res_x(a, xin, y, x - 1, addv), e = res_x(a, xin, y, x + 1, addv);
      double s = res_x(a, xin, y - 1, x, addv), n = res_x(a, xin, y + 1, x, addv);
      double r = g - (L.d * xc - L.cx * (w + e) - L.cy * (s + n));
      xout[k] = xc;
      r32 = (float)r;
      gg = g * g;
      rr = r * r;
      rs = r;
      gs = g;
    }
    a.r32[k] = r32;
  }
  double bg = block_sum<BS>(gg, sh);
  double br = block_sum<BS>(rr, sh);
  double bs = block_sum<BS>(rs, sh);
  double bgs = block_sum<BS>(gs, sh);
  const int nb = gridDim.x;
  if (threadIdx.x == 0) {
    store_partial(a.partials + blockIdx.x, bg);
    store_partial(a.partials + nb + blockIdx.x, br);
    store_partial(a.partials + 2 * nb + blockIdx.x, bs);
    store_partial(a.partials + 3 * nb + blockIdx.x, bgs);
  }
  if (!last_block_arrive<BS>(a.counter, &flag)) return;
  double tg = reduce_partials<BS>(a.partials, nb, sh, false);
  double tr = reduce_partials<BS>(a.partials + nb, nb, sh, false);
  double ts = reduce_partials<BS>(a.partials + 2 * nb, nb, sh, false);
  double tgs = reduce_partials<BS>(a.partials + 3 * nb, nb, sh, false);
  if (threadIdx.x == 0) {
    Ctl* o = a.outer;
    Ctl* c = a.inner;
    Ctl* in = a.inner;
    (void)in;
    // singular operator: norms of the zero-mean projections
    if (a.singular) {
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
    o->inner_total += c->iters;
    o->xcur = cur ^ 1;
    if (tg == 0.0 || rel <= a.rtol || !(rel == rel) || a.last) o->odone = 1;
    // the inner solve starts from x = 0 with g = r (its initial residual)
    c->bb = tr;
    c->rtol = a.inner_rtol;
    c->mean_r = a.singular ? ts / c->nwet : 0.0;
    c->done = 0;
    c->iters = 0;
    c->x_valid = 0;
    c->alpha = c->beta = c->mean_z = c->rz = 0.0;
    c->relres = 1.0;
  }
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

__global__ void k_reset_outer(Ctl* o, double nwet, double rtol) {
  if (threadIdx.x != 0) return;
  o->odone = 0;
  o->orefine = 0;
  o->inner_total = 0;
  o->obb = o->orr = 0.0;
  o->orelres = 1.0;
  o->nwet = nwet;
  o->rtol = rtol;
}

// mean of the current FP64 iterate (singular operator)
__global__ __launch_bounds__(BS) void k_mean64(LevelDev L, const double* x64a,
                                               const double* x64b,
                                               const Ctl* outer,
                                               double* partials,
                                               unsigned* counter, double nwet,
                                               double* out) {
  __shared__ double sh[BS];
  __shared__ int flag;
  const double* x = outer->xcur ? x64b : x64a;
  int64_t p = (int64_t)blockIdx.x * BS + threadIdx.x;
  double s = 0.0;
  if (p < (int64_t)L.ny * L.nx) {
    int y = (int)(p / L.nx), xx = (int)(p - (int64_t)y * L.nx);
    int64_t k = li(L, y, xx);
    if (L.m[k]) s = x[k];
  }
  double b = block_sum<BS>(s, sh);
  if (threadIdx.x == 0) store_partial(partials + blockIdx.x, b);
  if (!last_block_arrive<BS>(counter, &flag)) return;
  double tot = reduce_partials<BS>(partials, gridDim.x, sh, false);
  if (threadIdx.x == 0) *out = tot / nwet;
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
__global__ __launch_bounds__(BS) void k_mean(LevelDev L, const T* x,
                                             double* partials,
                                             unsigned* counter, double nwet,
                                             double* out) {
  __shared__ double sh[BS];
  __shared__ int flag;
  int64_t p = (int64_t)blockIdx.x * BS + threadIdx.x;
  double s = 0.0;
  if (p < (int64_t)L.ny * L.nx) {
    int y = (int)(p / L.nx), xx = (int)(p - (int64_t)y * L.nx);
    int64_t k = li(L, y, xx);
    if (L.m[k]) s = (double)x[k];
  }
  double b = block_sum<BS>(s, sh);
  if (threadIdx.x == 0) store_partial(partials + blockIdx.x, b);
  if (!last_block_arrive<BS>(counter, &flag)) return;
  double tot = reduce_partials<BS>(partials, gridDim.x, sh, false);
  if (threadIdx.x == 0) *out = tot / nwet;
}

template <class T>
__global__ void k_sub_mean(LevelDev L, T* x, const double* mean) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)L.ny * L.nx) return;
  int y = (int)(p / L.nx), xx = (int)(p - (int64_t)y * L.nx);
  int64_t k = li(L, y, xx);
  if (L.m[k]) x[k] = x[k] - T(*mean);
}

template <class T>
__global__ void k_stats(Ctl* c, const Ctl* outer_or_null) {
  if (threadIdx.x != 0) return;
  if (outer_or_null) {
    const Ctl* o = outer_or_null;
    c->stats[0] = (double)o->inner_total;
    c->stats[1] = o->orelres;
    c->stats[2] = (o->orelres <= c->rtol || o->obb == 0.0) ? 1.0 : 0.0;
    c->stats[3] = (double)o->orefine;
  } else {
    c->stats[0] = (double)c->iters;
    c->stats[1] = c->relres;
    c->stats[2] = (c->bb == 0.0 || c->relres <= c->rtol) ? 1.0 : 0.0;
    c->stats[3] = 0.0;
  }
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
  double* partials = nullptr;
  unsigned* counters = nullptr;  // 8 counters
  double* scal = nullptr;        // device scalars
  f2d::Ctl* outer = nullptr;
  int have_x64 = 0;
  int last_iters = 0;
  double inner_rtol = 1e-4;
  std::vector<void*> allocs;
};

namespace f2d {

static void* dalloc(f2d_poisson* h, size_t bytes) {
  void* p = nullptr;
  F2D_HIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
  F2D_HIP(hipMemset(p, 0, bytes < 256 ? 256 : bytes));
  h->allocs.push_back(p);
  return p;
}

template <class T>
static void build_pcg(f2d_poisson* h, Pcg<T>& P) {
  const int tail_points = h->opts.tail_points > 0 ? h->opts.tail_points : 4096;
  // level 0 in the padded layout of the caller's arrays
  HostLevel l0;
  l0.d = h->base;
  l0.m = h->m0;
  l0.d.m = h->m0;
  l0.n = (size_t)(h->NYP * h->NXP);
  P.lv.push_back(l0);
  // host copy of the level-0 mask to derive the coarse masks
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
    int cny = ny / 2, cnx = nx / 2;
    std::vector<int8_t> mc((size_t)cny * cnx);
    for (int J = 0; J < cny; J++)
      for (int I = 0; I < cnx; I++) mc[(size_t)J * cnx + I] = mprev[(size_t)(2 * J) * nx + 2 * I];
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
    P.lv.push_back(c);
    mprev.swap(mc);
    ny = cny;
    nx = cnx;
  }
  P.L = (int)P.lv.size();
  // tail start: first coarse level small enough for one workgroup's LDS
  P.Lt = P.L;
  for (int l = 1; l < P.L; l++) {
    size_t pts = (size_t)P.lv[l].d.ny * P.lv[l].d.nx;
    if ((int)pts > tail_points) continue;
    size_t bytes = 0;
    for (int k = l; k < P.L; k++) {
      size_t n = (size_t)P.lv[k].d.ny * P.lv[k].d.nx;
      bytes += 2 * ((n * sizeof(T) + 15) / 16) * 16 + ((n + 15) / 16) * 16;
    }
    if (bytes <= 150 * 1024 && P.L - l <= MAXL) {
      P.Lt = l;
      P.tail_lds = bytes;
      break;
    }
  }
  // vectors
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
    for (int k = 0; k < ta.nl; k++) {
      const HostLevel& v = P.lv[P.Lt + k];
      ta.lv[k] = v.d;
      size_t n = (size_t)v.d.ny * v.d.nx;
      ta.off_r[k] = off;
      off += (int)(((n * sizeof(T) + 15) / 16) * 16);
      ta.off_x[k] = off;
      off += (int)(((n * sizeof(T) + 15) / 16) * 16);
      ta.off_m[k] = off;
      off += (int)(((n + 15) / 16) * 16);
    }
    P.tail_lds = off;
    const LevelDev& cl = ta.lv[ta.nl - 1];
    int cs = h->opts.coarse_sweeps;
    if (cs <= 0) cs = 2 * (cl.ny > cl.nx ? cl.ny : cl.nx) + 2;
    ta.coarse_sweeps = cs;
    ta.r_in = (const T*)P.lv[P.Lt].r;
    ta.e_out = (T*)P.lv[P.Lt].e;
    ta.ctl = P.ctl;
  }
}

template <class T, int NU>
static void set_tail_attr(size_t bytes) {
  if (bytes > 64 * 1024)
    F2D_HIP(hipFuncSetAttribute((const void*)k_tail<T, NU>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
}

template <class T, int NU>
static void launch_down(const DownArgs<T>& a, bool cg, bool coarse, dim3 g,
                        hipStream_t s) {
  if (cg) {
    if (coarse) k_down<T, NU, true, true><<<g, BS, 0, s>>>(a);
    else k_down<T, NU, true, false><<<g, BS, 0, s>>>(a);
  } else {
    if (coarse) k_down<T, NU, false, true><<<g, BS, 0, s>>>(a);
    else k_down<T, NU, false, false><<<g, BS, 0, s>>>(a);
  }
}

template <class T, int NU>
static void launch_up(const UpArgs<T>& a, bool fine, bool coarse, dim3 g,
                      hipStream_t s) {
  if (fine) {
    if (coarse) k_up<T, NU, true, true><<<g, BS, 0, s>>>(a);
    else k_up<T, NU, true, false><<<g, BS, 0, s>>>(a);
  } else {
    if (coarse) k_up<T, NU, false, true><<<g, BS, 0, s>>>(a);
    else k_up<T, NU, false, false><<<g, BS, 0, s>>>(a);
  }
}

static dim3 tiles32(const LevelDev& L) {
  return dim3((unsigned)((L.nx + 31) / 32), (unsigned)((L.ny + 31) / 32));
}

// One PCG iteration k (k = 0: preconditioning of the initial residual).
template <class T, int NU>
static void pcg_iteration(f2d_poisson* h, Pcg<T>& P, int k, T* x,
                          const Ctl* outer, bool inner, bool last,
                          hipStream_t s) {
  const LevelDev& F = P.lv[0].d;
  const bool has_coarse = P.L > 1;
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
    a.partials = h->partials;
    a.counter = h->counters + 0;
    a.first = (k == 0);
    a.iter = k;
    a.inner = inner;
    launch_down<T, NU>(a, true, has_coarse, tiles32(F), s);
  }
  if (last) return;
  // coarse levels down
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
    launch_down<T, NU>(a, false, c, tiles32(a.F), s);
  }
  if (P.Lt < P.L) {
    TailArgs<T> ta = P.tail;
    ta.outer = outer;
    k_tail<T, NU><<<1, TAIL_BS, P.tail_lds, s>>>(ta);
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
    launch_up<T, NU>(a, false, c, tiles32(a.F), s);
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
    a.partials = h->partials;
    a.counter = h->counters + 1;
    a.first = (k == 0);
    a.singular = h->singular;
    launch_up<T, NU>(a, true, has_coarse, tiles32(F), s);
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
    a.partials = h->partials;
    a.counter = h->counters + 2;
    a.first = (k == 0);
    int nb = grid1d((int64_t)F.ny * F.nx, BS);
    k_pupdate<T><<<nb, BS, 0, s>>>(a);
  }
}

template <class T, int NU>
static void pcg_loop(f2d_poisson* h, Pcg<T>& P, T* x, const Ctl* outer,
                     bool inner, int maxit, bool sync_checks, hipStream_t s) {
  int* pinned = nullptr;
  int est = h->last_iters > 0 ? h->last_iters : 1;
  for (int k = 0; k <= maxit; k++) {
    pcg_iteration<T, NU>(h, P, k, x, outer, inner, k == maxit, s);
    if (sync_checks && k + 1 >= est && k < maxit) {
      int done = 0;
      F2D_HIP(hipMemcpyAsync(&done, &P.ctl->done, sizeof(int),
                             hipMemcpyDeviceToHost, s));
      F2D_HIP(hipStreamSynchronize(s));
      if (done) break;
    }
  }
  (void)pinned;
}

template <class T, class TB, int NU>
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
  ia.partials = h->partials;
  ia.counter = h->counters + 3;
  ia.singular = h->singular;
  ia.rtol = rtol;
  int nb = grid1d((int64_t)F.ny * F.nx, BS);
  if (gmode == G_FROM_B)
    k_init<T, TB, G_FROM_B><<<nb, BS, 0, s>>>(ia);
  else
    k_init<T, TB, G_FROM_C2V><<<nb, BS, 0, s>>>(ia);
  pcg_loop<T, NU>(h, P, x, nullptr, false, maxit, sync, s);
  if (h->singular) {
    k_mean<T><<<nb, BS, 0, s>>>(F, x, h->partials, h->counters + 4, h->nwet,
                                 h->scal);
    k_sub_mean<T><<<nb, BS, 0, s>>>(F, x, h->scal);
  }
  k_stats<T><<<1, 64, 0, s>>>(P.ctl, nullptr);
}

template <class TB, class TX, int NU>
static void solve_refine(f2d_poisson* h, const TB* b, TX* x, int gmode,
                         double rtol, int maxit, bool sync, hipStream_t s) {
  Pcg<float>& P = h->p32;
  const LevelDev& F = P.lv[0].d;
  int nb = grid1d((int64_t)F.ny * F.nx, BS);
  if (!h->have_x64) {
    // seed the FP64 iterate from the caller's array (buffer 0 is current)
    k_refine_seed<TX><<<nb, BS, 0, s>>>(F, x, h->x64[0]);
    h->have_x64 = 1;
  }
  k_reset_outer<<<1, 64, 0, s>>>(h->outer, h->nwet, rtol);
  if (h->singular)
    k_gmean<TB><<<nb, BS, 0, s>>>(F, b, gmode, h->partials, h->counters + 6,
                                  h->nwet, h->scal + 1);
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
    ra.partials = h->partials;
    ra.counter = h->counters + 5;
    ra.gmode = gmode;
    ra.singular = h->singular;
    ra.add_e = it > 0;
    ra.last = it == nref;
    ra.rtol = rtol;
    ra.inner_rtol = h->inner_rtol;
    ra.gmean = h->scal + 1;
    k_refine_res<TB><<<nb, BS, 0, s>>>(ra);
    if (it == nref) break;
    // inner FP32 solve of K e = r32 from e = 0 (at most inner_it
    // iterations), predicated off once the outer loop has converged
    for (int k = 0; k <= inner_it; k++)
      pcg_iteration<float, NU>(h, P, k, h->e32, h->outer, true, k == inner_it,
                               s);
    if (sync && it >= 1) {
      int od = 0;
      F2D_HIP(hipMemcpyAsync(&od, &h->outer->odone, sizeof(int),
                             hipMemcpyDeviceToHost, s));
      F2D_HIP(hipStreamSynchronize(s));
      if (od) {
        // one more residual pass is not needed: the last k_refine_res set
        // odone after writing the current iterate
        break;
      }
    }
  }
  const double* shift = nullptr;
  if (h->singular) {
    k_mean64<<<nb, BS, 0, s>>>(F, h->x64[0], h->x64[1], h->outer, h->partials,
                               h->counters + 4, h->nwet, h->scal);
    shift = h->scal;
  }
  k_refine_out<TX><<<nb, BS, 0, s>>>(F, h->x64[0], h->x64[1], h->outer, shift,
                                     x);
  k_stats<float><<<1, 64, 0, s>>>(P.ctl, h->outer);
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
  const int nu = h->opts.nu;
  const int mode = h->opts.mode;
  if (mode == F2D_POISSON_F64) {
    if (nu == 1)
      solve_plain<double, TB, 1>(h, h->p64, b, (double*)x, gmode, rtol, maxit, sync, s);
    else
      solve_plain<double, TB, 2>(h, h->p64, b, (double*)x, gmode, rtol, maxit, sync, s);
    fill_x<double>(h, (double*)x, s);
  } else if (mode == F2D_POISSON_F32) {
    if (nu == 1)
      solve_plain<float, TB, 1>(h, h->p32, b, (float*)x, gmode, rtol, maxit, sync, s);
    else
      solve_plain<float, TB, 2>(h, h->p32, b, (float*)x, gmode, rtol, maxit, sync, s);
    fill_x<float>(h, (float*)x, s);
  } else {
    // x has the dtype of b
    if (nu == 1)
      solve_refine<TB, TB, 1>(h, b, (TB*)x, gmode, rtol, maxit, sync, s);
    else
      solve_refine<TB, TB, 2>(h, b, (TB*)x, gmode, rtol, maxit, sync, s);
    fill_x<TB>(h, (TB*)x, s);
  }
}

}  // namespace f2d

using namespace f2d;

extern "C" {

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
    h->opts.tail_points = 4096;
    if (opts) {
      h->opts = *opts;
      if (h->opts.nu <= 0) h->opts.nu = 2;
      if (h->opts.refine_inner <= 0) h->opts.refine_inner = 4;
      if (h->opts.tail_points <= 0) h->opts.tail_points = 4096;
    }
    F2D_REQUIRE(h->opts.nu == 1 || h->opts.nu == 2, "opts.nu must be 1 or 2");
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
    int nbmax = (int)((h->ny * h->nx + 255) / 256) + 64;
    h->partials = (double*)dalloc(h, sizeof(double) * 3 * nbmax);
    h->counters = (unsigned*)dalloc(h, sizeof(unsigned) * 8);
    h->scal = (double*)dalloc(h, sizeof(double) * 8);
    h->outer = (Ctl*)dalloc(h, sizeof(Ctl));
    if (h->opts.mode == F2D_POISSON_F64) {
      build_pcg<double>(h, h->p64);
      if (h->p64.Lt < h->p64.L) {
        set_tail_attr<double, 1>(h->p64.tail_lds);
        set_tail_attr<double, 2>(h->p64.tail_lds);
      }
    } else {
      build_pcg<float>(h, h->p32);
      if (h->p32.Lt < h->p32.L) {
        set_tail_attr<float, 1>(h->p32.tail_lds);
        set_tail_attr<float, 2>(h->p32.tail_lds);
      }
      if (h->opts.mode == F2D_POISSON_F32_REFINE) {
        h->x64[0] = (double*)dalloc(h, N * sizeof(double));
        h->x64[1] = (double*)dalloc(h, N * sizeof(double));
        h->e32 = (float*)dalloc(h, N * sizeof(float));
      }
    }
    F2D_HIP(hipDeviceSynchronize());
    *out = h;
  });
  if (rc != 0 && h) {
    for (void* p : h->allocs) (void)hipFree(p);
    delete h;
  }
  return rc;
}

int f2d_poisson_destroy(f2d_poisson* h) {
  return guarded([&] {
    if (!h) return;
    F2D_HIP(hipDeviceSynchronize());
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
    if (mode == F2D_POISSON_F64)
      do_solve<double>(h, (const double*)b, x, gmode, rtol, maxit, sync, s);
    else
      do_solve<float>(h, (const float*)b, x, gmode, rtol, maxit, sync, s);
    if (sync) {
      Ctl* c = mode == F2D_POISSON_F64 ? h->p64.ctl : h->p32.ctl;
      double st[4];
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

const double* f2d_poisson_stats(f2d_poisson* h) {
  if (!h) return nullptr;
  Ctl* c = h->opts.mode == F2D_POISSON_F64 ? h->p64.ctl : h->p32.ctl;
  return c->stats;
}

}  // extern "C"
