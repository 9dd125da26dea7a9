// f2d_neumann.hip -- Jacobi-preconditioned CG for the Neumann ("c") Poisson
// operator of get_Laplacian_sparse (elliptic.py:114-195): the pressure
// projection of the euler model (operators.py:114-119).
//
// K = -A is symmetric positive (semi)definite on the unknowns:
//   (K x)_P = d_P x_P - sum_nb c_nb x_nb,  d_P = sum_nb c_nb + maindiag,
// c = dy/dx for the E/W and dx/dy for the N/S neighbours that are unknowns
// (periodic wrap in x, and in y as an extension).  Vectors live on the padded
// grid, zero outside the unknowns, in FP64 whatever the caller's dtype.
// Reductions are per-block partial sums finished by one block in a fixed
// order (deterministic); alpha, beta and the convergence flag stay on the
// device, so an iteration is five launches and no host round trip.
#include "f2d_common.hpp"
#include "f2d_neumann.hpp"

#include <vector>

namespace f2d {

struct NeuArgs {
  int ny, nx, NX;  // interior sizes, padded row length
  int xper, yper;
  double cx, cy;
};

// scalars: 0 rz, 1 gg, 2 done, 3 iterations, 4 rr, 5 alpha, 6 beta, 7 maxit
struct NeumannSolver {
  NeuArgs a{};
  int64_t n = 0;  // padded points
  int nb = 0;     // reduction blocks
  int8_t* m = nullptr;
  double* dg = nullptr;
  double *x = nullptr, *g = nullptr, *r = nullptr, *z = nullptr, *p = nullptr,
         *q = nullptr;
  double* part = nullptr;  // partial sums (max(3, ncomp) * nb)
  double* sc = nullptr;    // scalars
  double* stats = nullptr;
  // singular operator (no mass term): the connected wet regions, each with
  // a constant null vector; the right-hand side is projected onto the range
  int ncomp = 0;
  int32_t* comp = nullptr;   // region of each unknown (-1 elsewhere)
  double* inv_cnt = nullptr; // 1 / unknowns per region
  double* cmean = nullptr;   // per-region mean of g
};

constexpr int NEU_MAX_COMP = 64;

constexpr int NEU_BS = 256;

__device__ __forceinline__ int64_t nbr_w(const NeuArgs& a, int j, int i) {
  const int iw = (i - 1 >= 3) ? i - 1 : (a.xper ? i - 1 + a.nx : i - 1);
  return (int64_t)j * a.NX + iw;
}
__device__ __forceinline__ int64_t nbr_e(const NeuArgs& a, int j, int i) {
  const int ie = (i + 1 < 3 + a.nx) ? i + 1 : (a.xper ? i + 1 - a.nx : i + 1);
  return (int64_t)j * a.NX + ie;
}
__device__ __forceinline__ int64_t nbr_s(const NeuArgs& a, int j, int i) {
  const int js = (j - 1 >= 3) ? j - 1 : (a.yper ? j - 1 + a.ny : j - 1);
  return (int64_t)js * a.NX + i;
}
__device__ __forceinline__ int64_t nbr_n(const NeuArgs& a, int j, int i) {
  const int jn = (j + 1 < 3 + a.ny) ? j + 1 : (a.yper ? j + 1 - a.ny : j + 1);
  return (int64_t)jn * a.NX + i;
}

// (K v) at unknown k = (j, i): neighbours outside the unknowns hold 0
__device__ __forceinline__ double kmul(const NeuArgs& a, const double* __restrict__ v,
                                       const double* __restrict__ dg, int j, int i,
                                       int64_t k) {
  const double ew = v[nbr_w(a, j, i)] + v[nbr_e(a, j, i)];
  const double sn = v[nbr_s(a, j, i)] + v[nbr_n(a, j, i)];
  return dg[k] * v[k] - (a.cx * ew + a.cy * sn);
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = NEU_BS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double fin_sum(const double* part, int nb, double* sh) {
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += NEU_BS) v += part[b];
  return block_sum(v, sh);
}

template <class T>
__global__ __launch_bounds__(NEU_BS) void k_neu_load(int64_t n, const int8_t* __restrict__ m,
                                                     const T* __restrict__ b,
                                                     const T* __restrict__ x,
                                                     double* __restrict__ g,
                                                     double* __restrict__ x64) {
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  if (k >= n) return;
  const bool u = m[k] != 0;
  g[k] = u ? -(double)b[k] : 0.0;
  x64[k] = u ? (double)x[k] : 0.0;
}

// r = g - K x, z = r / d; partials of (r, z) and (g, g)
__global__ __launch_bounds__(NEU_BS) void k_neu_res(NeuArgs a, int64_t n,
                                                    const int8_t* __restrict__ m,
                                                    const double* __restrict__ dg,
                                                    const double* __restrict__ g,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ r,
                                                    double* __restrict__ z,
                                                    double* __restrict__ part) {
  __shared__ double sh[NEU_BS];
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  double rz = 0.0, gg = 0.0;
  if (k < n) {
    double rv = 0.0, zv = 0.0;
    if (m[k]) {
      const int j = (int)(k / a.NX), i = (int)(k - (int64_t)j * a.NX);
      rv = g[k] - kmul(a, x, dg, j, i, k);
      zv = rv / dg[k];
      rz = rv * zv;
      gg = g[k] * g[k];
    }
    r[k] = rv;
    z[k] = zv;
  }
  const double rr = block_sum(k < n ? r[k] * r[k] : 0.0, sh);
  rz = block_sum(rz, sh);
  gg = block_sum(gg, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rz;
    part[gridDim.x + blockIdx.x] = gg;
    part[2 * gridDim.x + blockIdx.x] = rr;
  }
}

// the reference norm of the convergence test is max(|g|, |r0|): a warm
// start measures against the right-hand side, and a right-hand side that
// is zero (or rounding noise) against the initial residual
__global__ __launch_bounds__(NEU_BS) void k_neu_init_fin(const double* __restrict__ part, int nb,
                                                         double rtol, int maxit,
                                                         double* __restrict__ sc) {
  __shared__ double sh[NEU_BS];
  const double rz = fin_sum(part, nb, sh);
  const double gg = fin_sum(part + nb, nb, sh);
  const double rr = fin_sum(part + 2 * nb, nb, sh);
  if (threadIdx.x == 0) {
    const double ref = gg > rr ? gg : rr;
    sc[0] = rz;
    sc[1] = ref;
    sc[3] = 0.0;
    sc[4] = rr;
    sc[7] = (double)maxit;
    sc[2] = (ref == 0.0 || maxit == 0 || sqrt(rr) <= rtol * sqrt(ref)) ? 1.0 : 0.0;
  }
}

// per-region sums of g (block partials, one reduction per region)
__global__ __launch_bounds__(NEU_BS) void k_neu_csum(int64_t n, const int32_t* __restrict__ comp,
                                                     const double* __restrict__ g, int ncomp,
                                                     double* __restrict__ part) {
  __shared__ double sh[NEU_BS];
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  const int c = k < n ? comp[k] : -1;
  const double v = c >= 0 ? g[k] : 0.0;
  for (int q = 0; q < ncomp; q++) {
    const double sq = block_sum(c == q ? v : 0.0, sh);
    if (threadIdx.x == 0) part[(int64_t)q * gridDim.x + blockIdx.x] = sq;
  }
}

// per-region means, in a fixed order of the block partials
__global__ __launch_bounds__(NEU_BS) void k_neu_cmean(const double* __restrict__ part, int nb,
                                                      int ncomp,
                                                      const double* __restrict__ inv_cnt,
                                                      double* __restrict__ cmean) {
  __shared__ double sh[NEU_BS];
  for (int q = 0; q < ncomp; q++) {
    const double sq = fin_sum(part + (int64_t)q * nb, nb, sh);
    if (threadIdx.x == 0) cmean[q] = sq * inv_cnt[q];
  }
}

__global__ __launch_bounds__(NEU_BS) void k_neu_project(int64_t n, const int32_t* __restrict__ comp,
                                                        const double* __restrict__ cmean,
                                                        double* __restrict__ g) {
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  if (k >= n) return;
  const int c = comp[k];
  if (c >= 0) g[k] -= cmean[c];
}

__global__ __launch_bounds__(NEU_BS) void k_neu_copy(int64_t n, const double* __restrict__ s,
                                                     double* __restrict__ d) {
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  if (k < n) d[k] = s[k];
}

// q = K p; partials of (p, q)
__global__ __launch_bounds__(NEU_BS) void k_neu_matvec(NeuArgs a, int64_t n,
                                                       const int8_t* __restrict__ m,
                                                       const double* __restrict__ dg,
                                                       const double* __restrict__ p,
                                                       double* __restrict__ q,
                                                       double* __restrict__ part,
                                                       const double* __restrict__ sc) {
  __shared__ double sh[NEU_BS];
  if (sc[2] != 0.0) return;  // converged: uniform over the grid
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  double pq = 0.0;
  if (k < n) {
    double qv = 0.0;
    if (m[k]) {
      const int j = (int)(k / a.NX), i = (int)(k - (int64_t)j * a.NX);
      qv = kmul(a, p, dg, j, i, k);
      pq = p[k] * qv;
    }
    q[k] = qv;
  }
  pq = block_sum(pq, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

__global__ __launch_bounds__(NEU_BS) void k_neu_alpha(const double* __restrict__ part, int nb,
                                                      double* __restrict__ sc) {
  __shared__ double sh[NEU_BS];
  if (sc[2] != 0.0) return;
  const double pq = fin_sum(part, nb, sh);
  if (threadIdx.x == 0) sc[5] = pq > 0.0 ? sc[0] / pq : 0.0;
}

// x += alpha p, r -= alpha q, z = r / d; partials of (r, z) and (r, r)
__global__ __launch_bounds__(NEU_BS) void k_neu_update(int64_t n, const int8_t* __restrict__ m,
                                                       const double* __restrict__ dg,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ q,
                                                       double* __restrict__ x,
                                                       double* __restrict__ r,
                                                       double* __restrict__ z,
                                                       double* __restrict__ part,
                                                       const double* __restrict__ sc) {
  __shared__ double sh[NEU_BS];
  if (sc[2] != 0.0) return;
  const double al = sc[5];
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (k < n && m[k]) {
    x[k] += al * p[k];
    const double rv = r[k] - al * q[k];
    const double zv = rv / dg[k];
    r[k] = rv;
    z[k] = zv;
    rz = rv * zv;
    rr = rv * rv;
  }
  rz = block_sum(rz, sh);
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rz;
    part[gridDim.x + blockIdx.x] = rr;
  }
}

__global__ __launch_bounds__(NEU_BS) void k_neu_beta(const double* __restrict__ part, int nb,
                                                     double rtol, double* __restrict__ sc) {
  __shared__ double sh[NEU_BS];
  if (sc[2] != 0.0) return;
  const double rzn = fin_sum(part, nb, sh);
  const double rr = fin_sum(part + nb, nb, sh);
  if (threadIdx.x == 0) {
    sc[6] = sc[0] > 0.0 ? rzn / sc[0] : 0.0;
    sc[0] = rzn;
    sc[4] = rr;
    sc[3] += 1.0;
    if (sqrt(rr) <= rtol * sqrt(sc[1]) || sc[3] >= sc[7]) sc[2] = 1.0;
  }
}

__global__ __launch_bounds__(NEU_BS) void k_neu_pupd(int64_t n, const int8_t* __restrict__ m,
                                                     const double* __restrict__ z,
                                                     double* __restrict__ p,
                                                     const double* __restrict__ sc) {
  if (sc[2] != 0.0) return;
  const double be = sc[6];
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  if (k < n && m[k]) p[k] = z[k] + be * p[k];
}

// x (caller's dtype) = x64 at the unknowns, with the periodic halo images
template <class T>
__global__ __launch_bounds__(NEU_BS) void k_neu_store(NeuArgs a, int64_t n,
                                                      const int8_t* __restrict__ m,
                                                      const double* __restrict__ x64,
                                                      T* __restrict__ x) {
  const int64_t k = (int64_t)blockIdx.x * NEU_BS + threadIdx.x;
  if (k >= n || !m[k]) return;
  const int j = (int)(k / a.NX), i = (int)(k - (int64_t)j * a.NX);
  const T v = (T)x64[k];
  x[k] = v;
  const int64_t dxs = !a.xper ? 0 : (i < 6 ? a.nx : (i >= a.nx ? -a.nx : 0));
  const int64_t dys = !a.yper ? 0 : (j < 6 ? (int64_t)a.ny * a.NX
                                           : (j >= a.ny ? -(int64_t)a.ny * a.NX : 0));
  if (dxs) x[k + dxs] = v;
  if (dys) x[k + dys] = v;
  if (dxs && dys) x[k + dxs + dys] = v;
}

__global__ void k_neu_stats(const double* __restrict__ sc, double rtol,
                            double* __restrict__ st) {
  if (threadIdx.x != 0) return;
  const double rel = sc[1] > 0.0 ? sqrt(sc[4] / sc[1]) : 0.0;
  const double conv = rel <= rtol ? 1.0 : 0.0;
  st[0] = sc[3];
  st[1] = rel;
  st[2] = conv;
  st[3] = 0.0;
  st[4] += conv > 0.0 ? 0.0 : 1.0;
  if (sc[3] > st[5]) st[5] = sc[3];
  st[6] += 1.0;
}

__global__ void k_neu_reset(double* st) {
  if (threadIdx.x < 8) st[threadIdx.x] = 0.0;
}

static double* neu_alloc(size_t bytes) {
  void* p = nullptr;
  F2D_HIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
  F2D_HIP(hipMemset(p, 0, bytes < 256 ? 256 : bytes));
  return (double*)p;
}

NeumannSolver* neu_create(const int8_t* m_host, int64_t ny, int64_t nx, int xper, int yper,
                          double dx, double dy, double maindiag) {
  NeumannSolver* s = new NeumannSolver();
  const int64_t NY = ny + 6, NX = nx + 6, n = NY * NX;
  s->a = NeuArgs{(int)ny, (int)nx, (int)NX, xper, yper, dy / dx, dx / dy};
  s->n = n;
  s->nb = (int)((n + NEU_BS - 1) / NEU_BS);
  // the diagonal of K: the coefficients of the unknown neighbours, summed in
  // the reference's order (west, east, south, north), plus maindiag
  std::vector<double> dgh(n, 0.0);
  auto unk = [&](int64_t j, int64_t i) -> bool {
    if (i < 3) i = xper ? i + nx : i;
    if (i >= 3 + nx) i = xper ? i - nx : i;
    if (j < 3) j = yper ? j + ny : j;
    if (j >= 3 + ny) j = yper ? j - ny : j;
    if (i < 0 || i >= NX || j < 0 || j >= NY) return false;
    return m_host[j * NX + i] != 0;
  };
  for (int64_t j = 0; j < NY; j++)
    for (int64_t i = 0; i < NX; i++) {
      if (!m_host[j * NX + i]) continue;
      double sum = 0.0;
      if (unk(j, i - 1)) sum += s->a.cx;
      if (unk(j, i + 1)) sum += s->a.cx;
      if (unk(j - 1, i)) sum += s->a.cy;
      if (unk(j + 1, i)) sum += s->a.cy;
      dgh[j * NX + i] = sum + maindiag;
      F2D_REQUIRE(dgh[j * NX + i] > 0.0,
                  "Neumann operator: an isolated unknown (no neighbour, no mass term)");
    }
  if (maindiag == 0.0) {
    // connected wet regions (4-neighbour, with the periodic wrap)
    std::vector<int32_t> comp(n, -1);
    std::vector<double> cnt;
    std::vector<int64_t> stack;
    for (int64_t k0 = 0; k0 < n; k0++) {
      if (!m_host[k0] || comp[k0] >= 0) continue;
      const int c = (int)cnt.size();
      F2D_REQUIRE(c < NEU_MAX_COMP,
                  "Neumann operator: more than 64 disconnected wet regions");
      cnt.push_back(0.0);
      comp[k0] = c;
      stack.push_back(k0);
      while (!stack.empty()) {
        const int64_t k = stack.back();
        stack.pop_back();
        cnt[c] += 1.0;
        const int64_t j = k / NX, i = k - j * NX;
        const int64_t nbj[4] = {j, j, j - 1, j + 1}, nbi[4] = {i - 1, i + 1, i, i};
        for (int t = 0; t < 4; t++) {
          int64_t jj = nbj[t], ii = nbi[t];
          if (ii < 3) ii = xper ? ii + nx : ii;
          if (ii >= 3 + nx) ii = xper ? ii - nx : ii;
          if (jj < 3) jj = yper ? jj + ny : jj;
          if (jj >= 3 + ny) jj = yper ? jj - ny : jj;
          if (ii < 0 || ii >= NX || jj < 0 || jj >= NY) continue;
          const int64_t kk = jj * NX + ii;
          if (m_host[kk] && comp[kk] < 0) {
            comp[kk] = c;
            stack.push_back(kk);
          }
        }
      }
    }
    s->ncomp = (int)cnt.size();
    std::vector<double> inv(cnt.size());
    for (size_t c = 0; c < cnt.size(); c++) inv[c] = 1.0 / cnt[c];
    s->comp = (int32_t*)neu_alloc(sizeof(int32_t) * n);
    F2D_HIP(hipMemcpy(s->comp, comp.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    s->inv_cnt = neu_alloc(sizeof(double) * (cnt.size() + 1));
    F2D_HIP(hipMemcpy(s->inv_cnt, inv.data(), sizeof(double) * cnt.size(),
                      hipMemcpyHostToDevice));
    s->cmean = neu_alloc(sizeof(double) * (cnt.size() + 1));
  }
  s->m = (int8_t*)neu_alloc(n);
  F2D_HIP(hipMemcpy(s->m, m_host, n, hipMemcpyHostToDevice));
  s->dg = neu_alloc(sizeof(double) * n);
  F2D_HIP(hipMemcpy(s->dg, dgh.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  s->x = neu_alloc(sizeof(double) * n);
  s->g = neu_alloc(sizeof(double) * n);
  s->r = neu_alloc(sizeof(double) * n);
  s->z = neu_alloc(sizeof(double) * n);
  s->p = neu_alloc(sizeof(double) * n);
  s->q = neu_alloc(sizeof(double) * n);
  s->part = neu_alloc(sizeof(double) * (s->ncomp > 3 ? s->ncomp : 3) * s->nb);
  s->sc = neu_alloc(sizeof(double) * 8);
  s->stats = neu_alloc(sizeof(double) * 8);
  return s;
}

void neu_destroy(NeumannSolver* s) {
  if (!s) return;
  for (void* p : {(void*)s->m, (void*)s->dg, (void*)s->x, (void*)s->g, (void*)s->r,
                  (void*)s->z, (void*)s->p, (void*)s->q, (void*)s->part, (void*)s->sc,
                  (void*)s->stats, (void*)s->comp, (void*)s->inv_cnt, (void*)s->cmean})
    if (p) (void)hipFree(p);
  delete s;
}

double* neu_stats(NeumannSolver* s) { return s->stats; }

void neu_reset_stats(NeumannSolver* s, hipStream_t st) {
  k_neu_reset<<<1, 64, 0, st>>>(s->stats);
}

void neu_solve(NeumannSolver* s, int dtype, const void* b, void* x, double rtol, int maxit,
               bool sync, int* iters, double* relres, hipStream_t st) {
  const int64_t n = s->n;
  const int nb = s->nb;
  if (dtype == F2D_F64)
    k_neu_load<double><<<nb, NEU_BS, 0, st>>>(n, s->m, (const double*)b, (const double*)x,
                                              s->g, s->x);
  else
    k_neu_load<float><<<nb, NEU_BS, 0, st>>>(n, s->m, (const float*)b, (const float*)x,
                                             s->g, s->x);
  if (s->ncomp > 0) {
    k_neu_csum<<<nb, NEU_BS, 0, st>>>(n, s->comp, s->g, s->ncomp, s->part);
    k_neu_cmean<<<1, NEU_BS, 0, st>>>(s->part, nb, s->ncomp, s->inv_cnt, s->cmean);
    k_neu_project<<<nb, NEU_BS, 0, st>>>(n, s->comp, s->cmean, s->g);
  }
  k_neu_res<<<nb, NEU_BS, 0, st>>>(s->a, n, s->m, s->dg, s->g, s->x, s->r, s->z, s->part);
  k_neu_init_fin<<<1, NEU_BS, 0, st>>>(s->part, nb, rtol, maxit, s->sc);
  k_neu_copy<<<nb, NEU_BS, 0, st>>>(n, s->z, s->p);
  for (int it = 0; it < maxit; it++) {
    k_neu_matvec<<<nb, NEU_BS, 0, st>>>(s->a, n, s->m, s->dg, s->p, s->q, s->part, s->sc);
    k_neu_alpha<<<1, NEU_BS, 0, st>>>(s->part, nb, s->sc);
    k_neu_update<<<nb, NEU_BS, 0, st>>>(n, s->m, s->dg, s->p, s->q, s->x, s->r, s->z,
                                        s->part, s->sc);
    k_neu_beta<<<1, NEU_BS, 0, st>>>(s->part, nb, rtol, s->sc);
    k_neu_pupd<<<nb, NEU_BS, 0, st>>>(n, s->m, s->z, s->p, s->sc);
    if (sync && (it % 8 == 7)) {
      double done = 0.0;
      F2D_HIP(hipMemcpyAsync(&done, s->sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
      F2D_HIP(hipStreamSynchronize(st));
      if (done != 0.0) break;
    }
  }
  if (dtype == F2D_F64)
    k_neu_store<double><<<nb, NEU_BS, 0, st>>>(s->a, n, s->m, s->x, (double*)x);
  else
    k_neu_store<float><<<nb, NEU_BS, 0, st>>>(s->a, n, s->m, s->x, (float*)x);
  k_neu_stats<<<1, 64, 0, st>>>(s->sc, rtol, s->stats);
  if (sync) {
    double stv[8];
    F2D_HIP(hipMemcpyAsync(stv, s->stats, sizeof(stv), hipMemcpyDeviceToHost, st));
    F2D_HIP(hipStreamSynchronize(st));
    if (iters) *iters = (int)stv[0];
    if (relres) *relres = stv[1];
  }
}

}  // namespace f2d
