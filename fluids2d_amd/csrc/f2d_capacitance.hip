// f2d_capacitance.hip -- masked Poisson/Helmholtz on the vertex grid as a
// direct solve: two spectral solves of the rectangle operator around a dense
// capacitance correction (see f2d_capacitance.hpp for the method).
//
// Replaces, for masks inside a rectangle the spectral solve handles, the
// factorisation and back-substitution of Poisson2D (elliptic.py:71-87) on
// the masked operator of get_Laplacian_sparse (elliptic.py:114-195).
//
// Kernels of one FP64 solve:
//   k_cap_rhs    b masked to the unknowns W (b_W), zero elsewhere
//   (3 spectral passes)   y = A_R^-1 b_W
//   k_cap_symv   partial products of the symmetric M = (-C)^-1 with y_B,
//                one 64 x 64 tile of its lower triangle per workgroup
//   k_cap_symv_sum  s = M y_B from the partials, written into b_W at B
//   (3 spectral passes)   y = A_R^-1 (b_W + P_B^T s)
//   k_cap_out    x = y at the unknowns W
// The FP32 mode wraps the same sequence, in FP32 with an FP32 M, into FP64
// refinement (k_cap_res forms the FP64 residual, k_cap_acc accumulates).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstring>
#include <vector>

#include "f2d_capacitance.hpp"
#include "f2d_fft.hpp"

namespace f2d {

// the padded geometry every kernel of this file takes
struct CapGeo {
  int ny, nx, NXP, NYP;
  int xper, yper;
  double cx, cy, d;  // (A x)_P = cx (x_E + x_W) + cy (x_N + x_S) - d x_P
};

__device__ __forceinline__ int wrap_i(const CapGeo& g, int i) {
  if (!g.xper) return i;
  return i < 3 ? i + g.nx : (i >= g.nx + 3 ? i - g.nx : i);
}
__device__ __forceinline__ int wrap_j(const CapGeo& g, int j) {
  if (!g.yper) return j;
  return j < 3 ? j + g.ny : (j >= g.ny + 3 ? j - g.ny : j);
}

// value of b at padded point k: given, or centerstovertices of the cell
// field (operators.py:126-133, the reference's summation order)
template <class TB>
__device__ __forceinline__ TB bval(const TB* b, int64_t k, int NXP, int gmode) {
  if (!gmode) return b[k];
  return TB(0.25) * (((b[k - NXP - 1] + b[k - 1]) + b[k - NXP]) + b[k]);
}

// rhs = b at the unknowns, 0 elsewhere (the points B included)
template <class TB, class T>
__global__ __launch_bounds__(256) void k_cap_rhs(CapGeo g, const TB* __restrict__ b,
                                                 int gmode, const int8_t* __restrict__ mw,
                                                 T* __restrict__ rhs) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= (int64_t)g.NYP * g.NXP) return;
  rhs[k] = mw[k] ? T(bval(b, k, g.NXP, gmode)) : T(0);
}

// x at an unknown neighbour (periodic wrap), 0 at a dry one
template <class TX>
__device__ __forceinline__ double xw(const CapGeo& g, const TX* x,
                                     const int8_t* mw, int j, int i) {
  const int64_t k = (int64_t)wrap_j(g, j) * g.NXP + wrap_i(g, i);
  return mw[k] ? (double)x[k] : 0.0;
}

// FP64 residual of the masked operator, r = b - A x at the unknowns (x read
// at unknowns only; xin NULL: x = 0), stored as T (0 elsewhere)
template <class TB, class TX, class T>
__device__ __forceinline__ void cap_res_at(const CapGeo& g, const TB* __restrict__ b,
                                           int gmode, const int8_t* __restrict__ mw,
                                           const TX* __restrict__ xin, T* __restrict__ r,
                                           int64_t k) {
  double v = 0.0;
  if (mw[k]) {
    v = (double)bval(b, k, g.NXP, gmode);
    if (xin) {
      const int j = (int)(k / g.NXP), i = (int)(k - (int64_t)j * g.NXP);
      const double ax = g.cx * (xw(g, xin, mw, j, i - 1) + xw(g, xin, mw, j, i + 1)) +
                        g.cy * (xw(g, xin, mw, j - 1, i) + xw(g, xin, mw, j + 1, i)) -
                        g.d * (double)xin[k];
      v -= ax;
    }
  }
  r[k] = T(v);
}

template <class TB, class TX, class T>
__global__ __launch_bounds__(256) void k_cap_res(CapGeo g, const TB* __restrict__ b,
                                                 int gmode, const int8_t* __restrict__ mw,
                                                 const TX* __restrict__ xin,
                                                 T* __restrict__ r) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= (int64_t)g.NYP * g.NXP) return;
  cap_res_at<TB, TX, T>(g, b, gmode, mw, xin, r, k);
}

// M = (-C)^-1 is symmetric: only the tiles (I, J), I >= J, of its lower
// triangle are stored, 64 x 64 entries each, row-major, in the order
// t = I (I + 1) / 2 + J, zero beyond n (diagonal tiles hold both
// triangles).  s = M y_B in two passes: each tile's workgroup forms
// M_IJ y_J (row partials) and, off the diagonal, M_IJ^T y_I (column
// partials); then each row of s sums its nt partials in a fixed order (FP64,
// deterministic).  Half the bytes of the full matrix are read.
constexpr int CT = 64;  // tile size

__device__ __forceinline__ void tile_ij(int t, int& I, int& J) {
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > t) i--;
  while ((i + 1) * (i + 2) / 2 <= t) i++;
  I = i;
  J = t - i * (i + 1) / 2;
}

template <class T>
__device__ __forceinline__ void load4(const T* p, double m[4]) {
  if (sizeof(T) == 4) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    m[0] = v.x;
    m[1] = v.y;
    m[2] = v.z;
    m[3] = v.w;
  } else {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    m[0] = a.x;
    m[1] = a.y;
    m[2] = b.x;
    m[3] = b.y;
  }
}

// partials of tile t (256 threads): part[t][0][r] = sum_c M_IJ[r][c] y_J[c],
// part[t][1][c] = sum_r M_IJ[r][c] y_I[r] (I > J).  Thread k holds columns
// 4 (k % 16) .. + 3 of rows k / 16 + 16 q, q = 0 .. 3.
template <class T>
__device__ __forceinline__ void cap_symv_tile(const T* __restrict__ Mt, int n,
                                              const int* __restrict__ bidx,
                                              const T* __restrict__ y,
                                              double* __restrict__ part, int t) {
  __shared__ double yi[CT], yj[CT], cs[4][CT];
  int I, J;
  tile_ij(t, I, J);
  const int tid = threadIdx.x;
  if (tid < CT) {
    const int gi = I * CT + tid;
    yi[tid] = gi < n ? (double)y[bidx[gi]] : 0.0;
  } else if (tid < 2 * CT) {
    const int gj = J * CT + tid - CT;
    yj[tid - CT] = gj < n ? (double)y[bidx[gj]] : 0.0;
  }
  __syncthreads();
  const int c4 = (tid & 15) * 4, r0 = tid >> 4;
  const T* tile = Mt + (int64_t)t * CT * CT;
  double m[4][4];
#pragma unroll
  for (int q = 0; q < 4; q++) load4<T>(tile + (r0 + 16 * q) * CT + c4, m[q]);
  double* pr = part + (int64_t)t * 2 * CT;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    double v = (m[q][0] * yj[c4] + m[q][1] * yj[c4 + 1]) +
               (m[q][2] * yj[c4 + 2] + m[q][3] * yj[c4 + 3]);
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    if ((tid & 15) == 0) pr[r0 + 16 * q] = v;
  }
  if (I == J) return;  // uniform per workgroup
  double cv[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    double v = (m[0][c] * yi[r0] + m[1][c] * yi[r0 + 16]) +
               (m[2][c] * yi[r0 + 32] + m[3][c] * yi[r0 + 48]);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    cv[c] = v;
  }
  const int wave = tid >> 6;
  if ((tid & 63) < 16) {
#pragma unroll
    for (int c = 0; c < 4; c++) cs[wave][c4 + c] = cv[c];
  }
  __syncthreads();
  if (tid < CT) pr[CT + tid] = (cs[0][tid] + cs[1][tid]) + (cs[2][tid] + cs[3][tid]);
}

template <class T>
__global__ __launch_bounds__(256) void k_cap_symv(const T* __restrict__ Mt, int n,
                                                  const int* __restrict__ bidx,
                                                  const T* __restrict__ y,
                                                  double* __restrict__ part) {
  cap_symv_tile<T>(Mt, n, bidx, y, part, (int)blockIdx.x);
}

// s_i = sum_{J <= I} part[(I, J)][0][r] + sum_{K > I} part[(K, I)][1][r],
// i = 64 I + r, into rhs at the point B_i
template <class T>
__device__ __forceinline__ void cap_symv_sum_at(int n, int nt, const double* __restrict__ part,
                                                const int* __restrict__ bidx,
                                                T* __restrict__ rhs, int i) {
  const int I = i / CT, r = i - I * CT;
  double acc = 0.0;
  for (int J = 0; J <= I; J++) acc += part[((int64_t)I * (I + 1) / 2 + J) * 2 * CT + r];
  for (int K = I + 1; K < nt; K++)
    acc += part[((int64_t)K * (K + 1) / 2 + I) * 2 * CT + CT + r];
  rhs[bidx[i]] = T(acc);
}

template <class T>
__global__ __launch_bounds__(256) void k_cap_symv_sum(int n, int nt,
                                                      const double* __restrict__ part,
                                                      const int* __restrict__ bidx,
                                                      T* __restrict__ rhs) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) cap_symv_sum_at<T>(n, nt, part, bidx, rhs, i);
}

// x = y at the unknowns
template <class T>
__global__ __launch_bounds__(256) void k_cap_out(int64_t n, const int8_t* __restrict__ mw,
                                                 const T* __restrict__ y, T* __restrict__ x) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < n && mw[k]) x[k] = y[k];
}

// x64 = x0 + e at the unknowns (x0 NULL: 0; x0 may be x64 itself)
template <class TX>
__global__ __launch_bounds__(256) void k_cap_acc(int64_t n, const int8_t* __restrict__ mw,
                                                 const TX* x0, const float* __restrict__ e,
                                                 double* x64) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < n && mw[k]) x64[k] = (x0 ? (double)x0[k] : 0.0) + (double)e[k];
}

__global__ __launch_bounds__(256) void k_cap_round(int64_t n, const int8_t* __restrict__ mw,
                                                   const double* __restrict__ x64,
                                                   float* __restrict__ x) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < n && mw[k]) x[k] = (float)x64[k];
}

__device__ __forceinline__ double bsum(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// per-block sums of r^2 and b^2 over the unknowns, r = b - A x (FP64)
template <class TB, class TX>
__global__ __launch_bounds__(256) void k_cap_norms(CapGeo g, const TB* __restrict__ b,
                                                   int gmode, const int8_t* __restrict__ mw,
                                                   const TX* __restrict__ x,
                                                   double* __restrict__ part) {
  __shared__ double sh[4];
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (k < (int64_t)g.NYP * g.NXP && mw[k]) {
    const int j = (int)(k / g.NXP), i = (int)(k - (int64_t)j * g.NXP);
    const double bv = (double)bval(b, k, g.NXP, gmode);
    const double ax = g.cx * (xw(g, x, mw, j, i - 1) + xw(g, x, mw, j, i + 1)) +
                      g.cy * (xw(g, x, mw, j - 1, i) + xw(g, x, mw, j + 1, i)) -
                      g.d * (double)x[k];
    rr = (bv - ax) * (bv - ax);
    bb = bv * bv;
  }
  rr = bsum(rr, sh);
  bb = bsum(bb, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[gridDim.x + blockIdx.x] = bb;
  }
}

// one block: relative residual into the solve statistics
// [iterations, relres, converged, refinements | failed, max, solves]
// (a direct solve is accepted unless its residual is above 1e-6 or not finite)
__global__ __launch_bounds__(256) void k_cap_stats(const double* __restrict__ part, int nb,
                                                   int passes, double* __restrict__ st) {
  __shared__ double sh[4];
  double a = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    a += part[i];
    c += part[nb + i];
  }
  a = bsum(a, sh);
  c = bsum(c, sh);
  if (threadIdx.x != 0) return;
  const double rel = c > 0.0 ? sqrt(a / c) : 0.0;
  const double conv = (rel <= 1e-6) ? 1.0 : 0.0;
  st[0] = (double)passes;
  st[1] = rel;
  st[2] = conv;
  st[3] = (double)(passes - 1);
  st[4] += conv > 0.0 ? 0.0 : 1.0;
  if ((double)(passes - 1) > st[5]) st[5] = (double)(passes - 1);
  st[6] += 1.0;
}

// one member of a batched FP32 solve (device table, blockIdx.y = member)
struct CapMember {
  const float* b;      // right-hand side at the vertices (padded)
  float* x;            // solution: warm start in, result out
  const int8_t* mw;    // unknowns
  float* rhs;          // scratch: right-hand side of the spectral solves
  float* y;            // scratch: spectral solution
  double* x64;         // FP64 accumulator
  const float* M;      // (-C)^-1, lower-triangle tiles
  const int* bidx;     // padded index of each point of B
  double* part;        // tile partials
  int n, nt, ntl, pad;
};

// from: 0 x = 0 (cold first pass), 1 the caller's x, 2 the accumulator
__global__ __launch_bounds__(256) void k_cap_res_b(CapGeo g, const CapMember* __restrict__ t,
                                                   int from) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= (int64_t)g.NYP * g.NXP) return;
  const CapMember& m = t[blockIdx.y];
  if (from == 1)
    cap_res_at<float, float, float>(g, m.b, 0, m.mw, m.x, m.rhs, k);
  else
    cap_res_at<float, double, float>(g, m.b, 0, m.mw, from ? m.x64 : nullptr, m.rhs, k);
}

// grid x: the largest member's tile count (the others return early)
__global__ __launch_bounds__(256) void k_cap_symv_b(const CapMember* __restrict__ t) {
  const CapMember& m = t[blockIdx.y];
  if ((int)blockIdx.x >= m.ntl) return;
  cap_symv_tile<float>(m.M, m.n, m.bidx, m.y, m.part, (int)blockIdx.x);
}

__global__ __launch_bounds__(256) void k_cap_symv_sum_b(const CapMember* __restrict__ t) {
  const CapMember& m = t[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < m.n) cap_symv_sum_at<float>(m.n, m.nt, m.part, m.bidx, m.rhs, i);
}

// x64 = x0 + y at the unknowns (from as in k_cap_res_b)
__global__ __launch_bounds__(256) void k_cap_acc_b(int64_t n, const CapMember* __restrict__ t,
                                                   int from) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const CapMember& m = t[blockIdx.y];
  if (k >= n || !m.mw[k]) return;
  const double x0 = from == 1 ? (double)m.x[k] : (from == 2 ? m.x64[k] : 0.0);
  m.x64[k] = x0 + (double)m.y[k];
}

__global__ __launch_bounds__(256) void k_cap_round_b(int64_t n,
                                                     const CapMember* __restrict__ t) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const CapMember& m = t[blockIdx.y];
  if (k < n && m.mw[k]) m.x[k] = (float)m.x64[k];
}

// C[i, j] = (A_R^-1)(B_i, B_j) from the Green's function G of the periodic
// doubled grid (ny2 x nx2, padded pitch nx2 + 6): sum over the images of B_j
// in the closed directions (odd extension: the mirror image about the wall
// carries the sign -), column-major with leading dimension n
__global__ __launch_bounds__(256) void k_cap_build(const double* __restrict__ G, int ny2,
                                                   int nx2, int xper, int yper,
                                                   const int* __restrict__ bx,
                                                   const int* __restrict__ by, int n,
                                                   double* __restrict__ C) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
  const int P = nx2 + 6;
  double v = 0.0;
  for (int sy = 0; sy < (yper ? 1 : 2); sy++) {
    int dy = sy ? by[r] + by[c] : by[r] - by[c];
    dy %= ny2;
    if (dy < 0) dy += ny2;
    for (int sx = 0; sx < (xper ? 1 : 2); sx++) {
      int dx = sx ? bx[r] + bx[c] : bx[r] - bx[c];
      dx %= nx2;
      if (dx < 0) dx += nx2;
      const double gv = G[(int64_t)(dy + 3) * P + dx + 3];
      v += ((sx + sy) & 1) ? -gv : gv;
    }
  }
  C[e] = -v;  // -C: symmetric positive definite
}

// the lower-triangle tiles of M from the inverse held in the lower
// triangle of A (column-major, n x n); zero beyond n
template <class T>
__global__ __launch_bounds__(256) void k_cap_tiles(const double* __restrict__ A, int n,
                                                   int64_t ntl, T* __restrict__ Mt) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ntl * CT * CT) return;
  const int t = (int)(e / (CT * CT)), w = (int)(e - (int64_t)t * CT * CT);
  int I, J;
  tile_ij(t, I, J);
  const int gi = I * CT + w / CT, gj = J * CT + w % CT;
  double v = 0.0;
  if (gi < n && gj < n) v = gi >= gj ? A[gi + (int64_t)gj * n] : A[gj + (int64_t)gi * n];
  Mt[e] = T(v);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct CapSolver {
  CapGeo g{};
  bool f64 = true;
  int n = 0, nt = 0, ntl = 0;  // boundary points, tile rows, stored tiles
  int8_t* mw = nullptr;     // unknowns (padded)
  int* bidx = nullptr;      // padded index of each point of B
  void* M = nullptr;        // (-C)^-1, lower-triangle tiles (FP64 or FP32)
  double* tpart = nullptr;  // tile partials of the product, 2 x 64 per tile
  void* rhs = nullptr;      // padded scratch: right-hand side (T)
  void* y = nullptr;        // padded scratch: spectral solution (T)
  double* x64 = nullptr;    // FP32 mode: FP64 accumulator (padded)
  double* part = nullptr;   // residual-norm partials
  int nbres = 0;
  bool solved = false;      // a warm start needs an earlier solve
  FftSolver* fft = nullptr; // the rectangle's spectral solve
  std::vector<void*> allocs;
};

static void* calloc_dev(CapSolver* c, size_t bytes) {
  void* p = nullptr;
  F2D_HIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
  F2D_HIP(hipMemset(p, 0, bytes < 256 ? 256 : bytes));
  c->allocs.push_back(p);
  return p;
}

static bool pw2(int64_t v) { return v >= 8 && (v & (v - 1)) == 0; }

// rectangle unknown (padded coordinates): every vertex strictly inside
static inline bool rect_unk(int64_t j, int64_t i, int64_t ny, int64_t nx, int xper,
                            int yper) {
  const bool oy = yper ? (j >= 3 && j < ny + 3) : (j >= 4 && j <= ny + 2);
  const bool ox = xper ? (i >= 3 && i < nx + 3) : (i >= 4 && i <= nx + 2);
  return oy && ox;
}

static std::vector<int64_t> boundary_points(const int8_t* unk, int64_t ny, int64_t nx,
                                            int xper, int yper) {
  const int64_t NXP = nx + 6, NYP = ny + 6;
  auto wi = [&](int64_t i) { return xper ? (i < 3 ? i + nx : (i >= nx + 3 ? i - nx : i)) : i; };
  auto wj = [&](int64_t j) { return yper ? (j < 3 ? j + ny : (j >= ny + 3 ? j - ny : j)) : j; };
  std::vector<int64_t> out;
  for (int64_t j = 0; j < NYP; j++)
    for (int64_t i = 0; i < NXP; i++) {
      if (!rect_unk(j, i, ny, nx, xper, yper) || unk[j * NXP + i]) continue;
      const int64_t nb[4][2] = {{j, i - 1}, {j, i + 1}, {j - 1, i}, {j + 1, i}};
      bool adj = false;
      for (auto& q : nb) {
        const int64_t jj = wj(q[0]), ii = wi(q[1]);
        if (jj >= 0 && jj < NYP && ii >= 0 && ii < NXP && unk[jj * NXP + ii]) adj = true;
      }
      if (adj) out.push_back(j * NXP + i);
    }
  return out;
}

int64_t cap_boundary(const int8_t* unk, int64_t ny, int64_t nx, int xper, int yper,
                     double maindiag, bool f64_mode) {
  if (!pw2(nx) || !pw2(ny)) return -1;
  if (xper && yper && maindiag == 0.0) return -1;  // singular rectangle
  const int64_t nx2 = xper ? nx : 2 * nx, ny2 = yper ? ny : 2 * ny;
  // the Green's function is an FP64 spectral solve on the doubled grid
  if (!fft_fits(ny2, nx2, 1, 1, true) || nx2 > 4096 || ny2 > 4096) return -1;
  if (!fft_fits(ny, nx, xper, yper, f64_mode)) return -1;
  const int64_t NXP = nx + 6, NYP = ny + 6;
  int64_t nw = 0, nr = 0;
  for (int64_t j = 0; j < NYP; j++)
    for (int64_t i = 0; i < NXP; i++) {
      const bool r = rect_unk(j, i, ny, nx, xper, yper);
      if (unk[j * NXP + i]) {
        if (!r) return -1;
        nw++;
      }
      nr += r;
    }
  if (nw == 0 || nw == nr) return -1;
  return (int64_t)boundary_points(unk, ny, nx, xper, yper).size();
}

static void rocblas_check(rocblas_status st, const char* what) {
  if (st != rocblas_status_success)
    throw Error{std::string(what) + ": rocBLAS/rocSOLVER status " + std::to_string((int)st)};
}

CapSolver* cap_create(const int8_t* unk, int64_t ny, int64_t nx, double dx, double dy,
                      double maindiag, int xper, int yper, bool f64_mode) {
  F2D_REQUIRE(cap_boundary(unk, ny, nx, xper, yper, maindiag, f64_mode) > 0,
              "capacitance solve does not apply to this operator");
  CapSolver* c = new CapSolver();
  FftSolver* gfs = nullptr;
  rocblas_handle rb = nullptr;
  try {
    const int64_t NXP = nx + 6, NYP = ny + 6, N = NXP * NYP;
    CapGeo& g = c->g;
    g.ny = (int)ny;
    g.nx = (int)nx;
    g.NXP = (int)NXP;
    g.NYP = (int)NYP;
    g.xper = xper;
    g.yper = yper;
    g.cx = dy / dx;
    g.cy = dx / dy;
    g.d = 2.0 * (g.cx + g.cy) + maindiag;
    c->f64 = f64_mode;
    const std::vector<int64_t> B = boundary_points(unk, ny, nx, xper, yper);
    const int n = (int)B.size();
    c->n = n;
    c->nt = (n + CT - 1) / CT;
    c->ntl = c->nt * (c->nt + 1) / 2;
    c->mw = (int8_t*)calloc_dev(c, (size_t)N);
    F2D_HIP(hipMemcpy(c->mw, unk, (size_t)N, hipMemcpyHostToDevice));
    std::vector<int> bi(n), bxv(n), byv(n);
    for (int k = 0; k < n; k++) {
      bi[k] = (int)B[k];
      byv[k] = (int)(B[k] / NXP) - 3;  // local coordinates: the first wall
      bxv[k] = (int)(B[k] % NXP) - 3;  // line (closed) or column (periodic) at 0
    }
    c->bidx = (int*)calloc_dev(c, sizeof(int) * n);
    F2D_HIP(hipMemcpy(c->bidx, bi.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    int* bxd = (int*)calloc_dev(c, sizeof(int) * n);
    int* byd = (int*)calloc_dev(c, sizeof(int) * n);
    F2D_HIP(hipMemcpy(bxd, bxv.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    F2D_HIP(hipMemcpy(byd, byv.data(), sizeof(int) * n, hipMemcpyHostToDevice));

    // Green's function of the periodic doubled grid: the response to a unit
    // source at the origin (padded (3, 3)).  maindiag = 0 with a closed
    // direction: the zero-mean solution, whose constant the odd images cancel.
    const int64_t nx2 = xper ? nx : 2 * nx, ny2 = yper ? ny : 2 * ny;
    const int64_t N2 = (nx2 + 6) * (ny2 + 6);
    gfs = fft_create(ny2, nx2, dx, dy, maindiag, true, false, 1, 1);
    double* src = nullptr;
    double* G = nullptr;
    F2D_HIP(hipMalloc(&src, sizeof(double) * N2));
    F2D_HIP(hipMalloc(&G, sizeof(double) * N2));
    F2D_HIP(hipMemset(src, 0, sizeof(double) * N2));
    F2D_HIP(hipMemset(G, 0, sizeof(double) * N2));
    const double one = 1.0;
    F2D_HIP(hipMemcpy(src + 3 * (nx2 + 6) + 3, &one, sizeof(double), hipMemcpyHostToDevice));
    fft_solve_f64(gfs, src, 0, G, 0.0, false, nullptr, 0);
    double* A = nullptr;
    F2D_HIP(hipMalloc(&A, sizeof(double) * (size_t)n * n));
    const int64_t ne = (int64_t)n * n;
    k_cap_build<<<(unsigned)((ne + 255) / 256), 256>>>(G, (int)ny2, (int)nx2, xper, yper, bxd,
                                                      byd, n, A);
    F2D_HIP(hipGetLastError());
    F2D_HIP(hipDeviceSynchronize());
    (void)hipFree(src);
    (void)hipFree(G);
    fft_destroy(gfs);
    gfs = nullptr;

    // M = (-C)^-1 by Cholesky factorisation and inversion (lower triangle)
    rocblas_check(rocblas_create_handle(&rb), "rocblas_create_handle");
    rocblas_int* info = nullptr;
    F2D_HIP(hipMalloc(&info, sizeof(rocblas_int)));
    rocblas_check(rocsolver_dpotrf(rb, rocblas_fill_lower, n, A, n, info), "rocsolver_dpotrf");
    rocblas_int ih = 0;
    F2D_HIP(hipMemcpy(&ih, info, sizeof(ih), hipMemcpyDeviceToHost));
    if (ih != 0) {
      (void)hipFree(info);
      (void)hipFree(A);
      throw Error{"capacitance matrix is not definite (potrf info " + std::to_string(ih) + ")"};
    }
    rocblas_check(rocsolver_dpotri(rb, rocblas_fill_lower, n, A, n, info), "rocsolver_dpotri");
    F2D_HIP(hipMemcpy(&ih, info, sizeof(ih), hipMemcpyDeviceToHost));
    (void)hipFree(info);
    rocblas_destroy_handle(rb);
    rb = nullptr;
    if (ih != 0) {
      (void)hipFree(A);
      throw Error{"capacitance matrix inversion failed (potri info " + std::to_string(ih) + ")"};
    }
    const int64_t nm = (int64_t)c->ntl * CT * CT;
    const size_t tsz = f64_mode ? sizeof(double) : sizeof(float);
    c->M = calloc_dev(c, tsz * (size_t)nm);
    c->tpart = (double*)calloc_dev(c, sizeof(double) * 2 * CT * (size_t)c->ntl);
    if (f64_mode)
      k_cap_tiles<double><<<(unsigned)((nm + 255) / 256), 256>>>(A, n, c->ntl, (double*)c->M);
    else
      k_cap_tiles<float><<<(unsigned)((nm + 255) / 256), 256>>>(A, n, c->ntl, (float*)c->M);
    F2D_HIP(hipGetLastError());
    F2D_HIP(hipDeviceSynchronize());
    (void)hipFree(A);

    c->rhs = calloc_dev(c, tsz * (size_t)N);
    c->y = calloc_dev(c, tsz * (size_t)N);
    if (!f64_mode) c->x64 = (double*)calloc_dev(c, sizeof(double) * (size_t)N);
    c->nbres = (int)((N + 255) / 256);
    c->part = (double*)calloc_dev(c, sizeof(double) * 2 * c->nbres);
    c->fft = fft_create(ny, nx, dx, dy, maindiag, f64_mode, !f64_mode, xper, yper);
    F2D_HIP(hipDeviceSynchronize());
  } catch (...) {
    if (rb) rocblas_destroy_handle(rb);
    if (gfs) fft_destroy(gfs);
    cap_destroy(c);
    throw;
  }
  return c;
}

void cap_destroy(CapSolver* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  if (c->fft) fft_destroy(c->fft);
  for (void* p : c->allocs) (void)hipFree(p);
  delete c;
}

int64_t cap_size(const CapSolver* c) { return c ? c->n : 0; }

static unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

// the capacitance correction between the two spectral solves of rhs
template <class T>
static void gemv(CapSolver* c, hipStream_t s) {
  k_cap_symv<T><<<c->ntl, 256, 0, s>>>((const T*)c->M, c->n, c->bidx, (const T*)c->y,
                                       c->tpart);
  k_cap_symv_sum<T><<<blocks(c->n), 256, 0, s>>>(c->n, c->nt, c->tpart, c->bidx,
                                                  (T*)c->rhs);
}

template <class TB, class TX>
static void measure_res(CapSolver* c, const TB* b, int gmode, const TX* x, int passes,
                        double* stats, hipStream_t s) {
  k_cap_norms<TB, TX><<<c->nbres, 256, 0, s>>>(c->g, b, gmode, c->mw, x, c->part);
  k_cap_stats<<<1, 256, 0, s>>>(c->part, c->nbres, passes, stats);
}

void cap_solve_f64(CapSolver* c, const double* b, int gmode, double* x, bool measure,
                   double* stats, hipStream_t s) {
  F2D_REQUIRE(c->f64, "capacitance solver built for FP32");
  const int64_t N = (int64_t)c->g.NYP * c->g.NXP;
  double* rhs = (double*)c->rhs;
  double* y = (double*)c->y;
  k_cap_rhs<double, double><<<blocks(N), 256, 0, s>>>(c->g, b, gmode, c->mw, rhs);
  fft_solve_f64(c->fft, rhs, 0, y, 0.0, false, nullptr, s);
  gemv<double>(c, s);
  fft_solve_f64(c->fft, rhs, 0, y, 0.0, false, nullptr, s);
  k_cap_out<double><<<blocks(N), 256, 0, s>>>(N, c->mw, y, x);
  if (measure && stats) measure_res<double, double>(c, b, gmode, x, 1, stats, s);
  c->solved = true;
}

void cap_solve_f32(CapSolver* c, const float* b, int gmode, float* x, int nref, bool warm,
                   bool measure, double* stats, hipStream_t s) {
  F2D_REQUIRE(!c->f64, "capacitance solver built for FP64");
  const int64_t N = (int64_t)c->g.NYP * c->g.NXP;
  float* rhs = (float*)c->rhs;
  float* y = (float*)c->y;
  if (nref == 0) {  // a plain FP32 solve, no FP64 refinement
    // the first spectral solve reads b at W and leaves b_W in rhs; the
    // second writes its result to x at W
    FftMask m1, m2;
    m1.in = c->mw;
    m1.out = rhs;
    m2.res = c->mw;
    fft_solve_f32_masked(c->fft, b, gmode, y, m1, s);
    gemv<float>(c, s);
    fft_solve_f32_masked(c->fft, rhs, 0, x, m2, s);
    if (measure && stats) measure_res<float, float>(c, b, gmode, x, 1, stats, s);
    c->solved = true;
    return;
  }
  const bool from_x = warm && c->solved;
  if (from_x && nref == 1 && !(measure && stats)) {
    // one refinement from the caller's x, fused into the spectral passes:
    // the first forward row pass forms the FP64 residual of x at W (kept in
    // rhs), the second inverse row pass adds the correction to x at W
    FftMask m1, m2;
    m1.in = c->mw;
    m1.out = rhs;
    m1.xr = x;
    m2.res = c->mw;
    fft_solve_f32_masked(c->fft, b, gmode, y, m1, s);
    gemv<float>(c, s);
    fft_solve_f32_masked(c->fft, rhs, 0, x, m2, s, true);
    c->solved = true;
    return;
  }
  const int passes = from_x ? (nref > 0 ? nref : 1) : (nref > 0 ? nref : 1) + 1;
  for (int p = 0; p < passes; p++) {
    // residual of the current iterate: the caller's x (warm, first pass),
    // x64 (later passes) or 0 (cold, first pass)
    if (p == 0 && from_x)
      k_cap_res<float, float, float><<<blocks(N), 256, 0, s>>>(c->g, b, gmode, c->mw, x, rhs);
    else if (p == 0)
      k_cap_res<float, double, float><<<blocks(N), 256, 0, s>>>(c->g, b, gmode, c->mw,
                                                               (const double*)nullptr, rhs);
    else
      k_cap_res<float, double, float><<<blocks(N), 256, 0, s>>>(c->g, b, gmode, c->mw,
                                                               c->x64, rhs);
    fft_solve_f32_refine(c->fft, rhs, 0, y, 0, false, 0.0, false, nullptr, s);
    gemv<float>(c, s);
    fft_solve_f32_refine(c->fft, rhs, 0, y, 0, false, 0.0, false, nullptr, s);
    if (p == 0 && from_x)
      k_cap_acc<float><<<blocks(N), 256, 0, s>>>(N, c->mw, x, y, c->x64);
    else if (p == 0)
      k_cap_acc<double><<<blocks(N), 256, 0, s>>>(N, c->mw, (const double*)nullptr, y, c->x64);
    else
      k_cap_acc<double><<<blocks(N), 256, 0, s>>>(N, c->mw, c->x64, y, c->x64);
  }
  k_cap_round<<<blocks(N), 256, 0, s>>>(N, c->mw, c->x64, x);
  if (measure && stats) measure_res<float, double>(c, b, gmode, c->x64, passes, stats, s);
  c->solved = true;
}

// ---------------------------------------------------------------------------
// batched FP32 solves of same-size members
// ---------------------------------------------------------------------------
struct CapBatch {
  std::vector<CapSolver*> c;
  std::vector<const float*> b;
  std::vector<float*> x;
  CapMember* t = nullptr;      // device member table
  FftBatchPtrs* ft = nullptr;  // device: rhs -> y through each member's spectrum
  FftBatchPtrs* fa = nullptr;  // plain solves: b (masked, kept in rhs) -> y
  FftBatchPtrs* fb = nullptr;  // plain solves: rhs -> x at the unknowns
  FftBatchPtrs* fr = nullptr;  // one warm refinement: residual of x (kept in rhs) -> y
  int maxntl = 0, maxn = 0;
};

static bool same_geo(const CapGeo& a, const CapGeo& b) {
  return a.ny == b.ny && a.nx == b.nx && a.xper == b.xper && a.yper == b.yper &&
         a.cx == b.cx && a.cy == b.cy && a.d == b.d;
}

CapBatch* cap_batch_create(CapSolver* const* c, int nm, const float* const* b,
                           float* const* x) {
  F2D_REQUIRE(nm >= 1 && nm <= 65535, "batch of 1 .. 65535 members");
  CapBatch* bt = new CapBatch();
  try {
    std::vector<CapMember> t(nm);
    std::vector<FftBatchPtrs> ft(nm), fa(nm), fb(nm), fr(nm);
    for (int i = 0; i < nm; i++) {
      CapSolver* m = c[i];
      F2D_REQUIRE(m && !m->f64, "batched capacitance solves are FP32 solvers");
      F2D_REQUIRE(b[i] && x[i], "batch member without b or x");
      F2D_REQUIRE(same_geo(m->g, c[0]->g) && fft_same_plan(m->fft, c[0]->fft),
                  "batch members must share the rectangle operator");
      bt->c.push_back(m);
      bt->b.push_back(b[i]);
      bt->x.push_back(x[i]);
      t[i] = CapMember{b[i], x[i], m->mw, (float*)m->rhs, (float*)m->y, m->x64,
                       (const float*)m->M, m->bidx, m->tpart, m->n, m->nt, m->ntl, 0};
      ft[i] = FftBatchPtrs{m->rhs, fft_spectrum32(m->fft), m->y, FftMask()};
      fa[i] = FftBatchPtrs{b[i], fft_spectrum32(m->fft), m->y, FftMask()};
      fa[i].m.in = m->mw;
      fa[i].m.out = m->rhs;
      fb[i] = FftBatchPtrs{m->rhs, fft_spectrum32(m->fft), x[i], FftMask()};
      fb[i].m.res = m->mw;
      fr[i] = fa[i];
      fr[i].m.xr = x[i];
      if (m->ntl > bt->maxntl) bt->maxntl = m->ntl;
      if (m->n > bt->maxn) bt->maxn = m->n;
    }
    F2D_HIP(hipMalloc(&bt->t, sizeof(CapMember) * nm));
    F2D_HIP(hipMalloc(&bt->ft, sizeof(FftBatchPtrs) * nm));
    F2D_HIP(hipMalloc(&bt->fa, sizeof(FftBatchPtrs) * nm));
    F2D_HIP(hipMalloc(&bt->fb, sizeof(FftBatchPtrs) * nm));
    F2D_HIP(hipMalloc(&bt->fr, sizeof(FftBatchPtrs) * nm));
    F2D_HIP(hipMemcpy(bt->t, t.data(), sizeof(CapMember) * nm, hipMemcpyHostToDevice));
    F2D_HIP(hipMemcpy(bt->ft, ft.data(), sizeof(FftBatchPtrs) * nm, hipMemcpyHostToDevice));
    F2D_HIP(hipMemcpy(bt->fa, fa.data(), sizeof(FftBatchPtrs) * nm, hipMemcpyHostToDevice));
    F2D_HIP(hipMemcpy(bt->fb, fb.data(), sizeof(FftBatchPtrs) * nm, hipMemcpyHostToDevice));
    F2D_HIP(hipMemcpy(bt->fr, fr.data(), sizeof(FftBatchPtrs) * nm, hipMemcpyHostToDevice));
  } catch (...) {
    cap_batch_destroy(bt);
    throw;
  }
  return bt;
}

void cap_batch_destroy(CapBatch* bt) {
  if (!bt) return;
  (void)hipDeviceSynchronize();
  if (bt->t) (void)hipFree(bt->t);
  if (bt->ft) (void)hipFree(bt->ft);
  if (bt->fa) (void)hipFree(bt->fa);
  if (bt->fb) (void)hipFree(bt->fb);
  if (bt->fr) (void)hipFree(bt->fr);
  delete bt;
}

bool cap_batch_warm(const CapBatch* bt) {
  for (const CapSolver* m : bt->c)
    if (!m->solved) return false;
  return true;
}

void cap_batch_solve(CapBatch* bt, int nref, bool warm, hipStream_t s) {
  const int nm = (int)bt->c.size();
  const CapGeo& g = bt->c[0]->g;
  const int64_t N = (int64_t)g.NYP * g.NXP;
  if (nref == 0) {  // plain FP32 solves (cap_solve_f32 with nref 0)
    fft_solve_f32_batch(bt->c[0]->fft, nm, bt->fa, s);
    k_cap_symv_b<<<dim3((unsigned)bt->maxntl, (unsigned)nm), 256, 0, s>>>(bt->t);
    k_cap_symv_sum_b<<<dim3(blocks(bt->maxn), (unsigned)nm), 256, 0, s>>>(bt->t);
    fft_solve_f32_batch(bt->c[0]->fft, nm, bt->fb, s);
    for (CapSolver* m : bt->c) m->solved = true;
    return;
  }
  const bool from_x = warm && cap_batch_warm(bt);
  bool none = true;
  for (const CapSolver* m : bt->c) none = none && !m->solved;
  if (warm && !from_x && !none) {
    // some members warm, some cold: each member's own solve
    for (int i = 0; i < nm; i++)
      cap_solve_f32(bt->c[i], bt->b[i], 0, bt->x[i], nref, true, false, nullptr, s);
    return;
  }
  if (from_x && nref == 1) {  // cap_solve_f32's fused warm refinement
    fft_solve_f32_batch(bt->c[0]->fft, nm, bt->fr, s);
    k_cap_symv_b<<<dim3((unsigned)bt->maxntl, (unsigned)nm), 256, 0, s>>>(bt->t);
    k_cap_symv_sum_b<<<dim3(blocks(bt->maxn), (unsigned)nm), 256, 0, s>>>(bt->t);
    fft_solve_f32_batch(bt->c[0]->fft, nm, bt->fb, s, true);
    for (CapSolver* m : bt->c) m->solved = true;
    return;
  }
  const int passes = from_x ? (nref > 0 ? nref : 1) : (nref > 0 ? nref : 1) + 1;
  const dim3 gp((unsigned)blocks(N), (unsigned)nm);
  for (int p = 0; p < passes; p++) {
    const int from = p > 0 ? 2 : (from_x ? 1 : 0);
    k_cap_res_b<<<gp, 256, 0, s>>>(g, bt->t, from);
    fft_solve_f32_batch(bt->c[0]->fft, nm, bt->ft, s);
    k_cap_symv_b<<<dim3((unsigned)bt->maxntl, (unsigned)nm), 256, 0, s>>>(bt->t);
    k_cap_symv_sum_b<<<dim3(blocks(bt->maxn), (unsigned)nm), 256, 0, s>>>(bt->t);
    fft_solve_f32_batch(bt->c[0]->fft, nm, bt->ft, s);
    k_cap_acc_b<<<gp, 256, 0, s>>>(N, bt->t, from);
  }
  k_cap_round_b<<<gp, 256, 0, s>>>(N, bt->t);
  for (CapSolver* m : bt->c) m->solved = true;
}

}  // namespace f2d
