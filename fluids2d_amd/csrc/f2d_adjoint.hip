// f2d_adjoint.hip -- adjoint (reverse-mode derivative) of the Euler
// vorticity-streamfunction RK stage on a fully wet, doubly periodic grid.
//
// One RK stage evaluates R(s) = T(s, U(s)): U(s) = perpgrad(A^-1 c2v(s))
// (operators.py:126-149, elliptic.py:80-87, contravariant), a linear map, and
// T(q, U) = -div F with face fluxes F = U r(q) (weno.py:346-364), r the
// WENO-Z5 reconstruction of the upwind stencil (all faces have order 6 on a
// fully wet periodic grid).  Given the adjoint lam of R(s), this file
// computes R'(s)^T lam = T_q^T lam + U'^T T_U^T lam:
//
//   k_adj_faces        per face f: lamF = lam[f] - lam[f-1] (the two cells
//                      the face bounds), lamU[f] = lamF r_f, and the six
//                      stencil weights W[m][f] = U_f lamF dr_f/dq_{f+m-3};
//                      lam is formed on the fly from the RK combination of
//                      the stage outputs' adjoints
//   k_adj_gather_perp  per cell c: sum over the faces whose stencils hold c
//                      of W[m][c-(m-3)] in x and y (a gather: deterministic,
//                      no atomics), and perpgrad^T of lamU
//   (A^-1 is self-adjoint: the same solve as the forward one)
//   k_adj_combine      c2v^T of the solve, plus the gathered term, plus the
//                      adjoint carried from the next stage
//
// dr/dq of WENO-Z5 follows the chain rule through the candidate values,
// the smoothness indicators beta_k, tau5 = |beta1 - beta3| and the weights
// (sign(beta1 - beta3) at the kink, 0 when equal); the upwind selection is
// piecewise constant in U.  All arrays use the padded (ny+6) x (nx+6)
// layout; outputs are written on the interior; adjoint inputs are read on
// the interior (periodic neighbours wrapped), the forward state q and U
// need their periodic halos filled.
#include "f2d_common.hpp"

namespace f2d {

// r = weno5z(x0..x4) and dr[j] = dr/dx_j
template <class T>
__device__ __forceinline__ T weno5z_grad(const T x[5], T dr[5]) {
  const T eps = T(1e-16);
  const T qi1 = (T(1. / 3.) * x[0] - T(7. / 6.) * x[1]) + T(11. / 6.) * x[2];
  const T qi2 = (T(-1. / 6.) * x[1] + T(5. / 6.) * x[2]) + T(1. / 3.) * x[3];
  const T qi3 = (T(1. / 3.) * x[2] + T(5. / 6.) * x[3]) - T(1. / 6.) * x[4];
  const T k1 = T(13. / 12.), k2 = T(0.25);
  const T a1 = (x[0] - T(2) * x[1]) + x[2], b1 = (x[0] - T(4) * x[1]) + T(3) * x[2];
  const T a2 = (x[1] - T(2) * x[2]) + x[3], b2 = x[1] - x[3];
  const T a3 = (x[2] - T(2) * x[3]) + x[4], b3 = (T(3) * x[2] - T(4) * x[3]) + x[4];
  const T be1 = k1 * (a1 * a1) + k2 * (b1 * b1);
  const T be2 = k1 * (a2 * a2) + k2 * (b2 * b2);
  const T be3 = k1 * (a3 * a3) + k2 * (b3 * b3);
  const T d13 = be1 - be3;
  const T tau = fabs(d13);
  const T sg = d13 > T(0) ? T(1) : (d13 < T(0) ? T(-1) : T(0));
  const T e1 = be1 + eps, e2 = be2 + eps, e3 = be3 + eps;
  const T w1 = T(0.1) * (T(1) + tau / e1);
  const T w2 = T(0.6) * (T(1) + tau / e2);
  const T w3 = T(0.3) * (T(1) + tau / e3);
  const T W = (w1 + w2) + w3;
  const T r = ((w1 * qi1 + w2 * qi2) + w3 * qi3) / W;
  // linear coefficients of qi_k, a_k, b_k
  const T cq1[5] = {T(1. / 3.), T(-7. / 6.), T(11. / 6.), T(0), T(0)};
  const T cq2[5] = {T(0), T(-1. / 6.), T(5. / 6.), T(1. / 3.), T(0)};
  const T cq3[5] = {T(0), T(0), T(1. / 3.), T(5. / 6.), T(-1. / 6.)};
  const T ca1[5] = {T(1), T(-2), T(1), T(0), T(0)};
  const T cb1[5] = {T(1), T(-4), T(3), T(0), T(0)};
  const T ca2[5] = {T(0), T(1), T(-2), T(1), T(0)};
  const T cb2[5] = {T(0), T(1), T(0), T(-1), T(0)};
  const T ca3[5] = {T(0), T(0), T(1), T(-2), T(1)};
  const T cb3[5] = {T(0), T(0), T(3), T(-4), T(1)};
  const T iW = T(1) / W;
#pragma unroll
  for (int j = 0; j < 5; j++) {
    const T db1 = T(2) * (k1 * a1 * ca1[j] + k2 * b1 * cb1[j]);
    const T db2 = T(2) * (k1 * a2 * ca2[j] + k2 * b2 * cb2[j]);
    const T db3 = T(2) * (k1 * a3 * ca3[j] + k2 * b3 * cb3[j]);
    const T dt = sg * (db1 - db3);
    const T dw1 = T(0.1) * (dt / e1 - tau * db1 / (e1 * e1));
    const T dw2 = T(0.6) * (dt / e2 - tau * db2 / (e2 * e2));
    const T dw3 = T(0.3) * (dt / e3 - tau * db3 / (e3 * e3));
    const T num = ((w1 * cq1[j] + w2 * cq2[j]) + w3 * cq3[j]) +
                  ((dw1 * (qi1 - r) + dw2 * (qi2 - r)) + dw3 * (qi3 - r));
    dr[j] = num * iW;
  }
  return r;
}

struct AdjArgs {
  int ny, nx, NXP;
};

__device__ __forceinline__ int wrapi(int v, int n) {
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

// lam of one RK stage's tendency: the combination c0 l0 + c1 l1 + c2 l2 of
// the adjoints of the stage outputs (an absent term has a NULL array)
template <class T>
struct LamComb {
  const T* l0;
  const T* l1;
  const T* l2;
  T c0, c1, c2;
  __device__ __forceinline__ T at(int64_t k) const {
    T v = c0 * l0[k];
    if (l1) v = v + c1 * l1[k];
    if (l2) v = v + c2 * l2[k];
    return v;
  }
};

// x faces and y faces at interior point (y, x): lamU and the stencil weights.
// lam is read at interior points only (periodic neighbours wrapped); q, Ux,
// Uy need their periodic halos filled.
template <class T>
__global__ __launch_bounds__(256) void k_adj_faces(AdjArgs a, const T* __restrict__ q,
                                                   const T* __restrict__ Ux,
                                                   const T* __restrict__ Uy,
                                                   LamComb<T> lam,
                                                   T* __restrict__ lUx,
                                                   T* __restrict__ lUy,
                                                   T* __restrict__ Wx,
                                                   T* __restrict__ Wy,
                                                   int64_t plane) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)a.ny * a.nx) return;
  const int y = (int)(idx / a.nx), x = (int)(idx - (int64_t)y * a.nx);
  const int64_t P = a.NXP;
  const int64_t k = (int64_t)(y + 3) * P + x + 3;
  const T lc = lam.at(k);
  const T lw = lam.at((int64_t)(y + 3) * P + wrapi(x - 1, a.nx) + 3);
  const T ls = lam.at((int64_t)(wrapi(y - 1, a.ny) + 3) * P + x + 3);
  // x face between cells (y, x-1) and (y, x)
  {
    const T lf = lc - lw;
    const T u = Ux[k];
    T xs[5], dr[5];
    const bool pos = u > T(0);
#pragma unroll
    for (int m = 0; m < 5; m++) xs[m] = pos ? q[k - 3 + m] : q[k + 2 - m];
    const T r = weno5z_grad<T>(xs, dr);
    lUx[k] = lf * r;
    const T s = u * lf;
#pragma unroll
    for (int m = 0; m < 6; m++) {
      // slot m <-> cell offset m - 3; pos: inputs j = m (offsets -3..1),
      // else inputs j = 5 - m (offsets 2..-2)
      T v = T(0);
      if (pos) {
        if (m < 5) v = s * dr[m];
      } else {
        if (m >= 1) v = s * dr[5 - m];
      }
      Wx[m * plane + k] = v;
    }
  }
  // y face between cells (y-1, x) and (y, x)
  {
    const T lf = lc - ls;
    const T u = Uy[k];
    T xs[5], dr[5];
    const bool pos = u > T(0);
#pragma unroll
    for (int m = 0; m < 5; m++) xs[m] = pos ? q[k + (m - 3) * P] : q[k + (2 - m) * P];
    const T r = weno5z_grad<T>(xs, dr);
    lUy[k] = lf * r;
    const T s = u * lf;
#pragma unroll
    for (int m = 0; m < 6; m++) {
      T v = T(0);
      if (pos) {
        if (m < 5) v = s * dr[m];
      } else {
        if (m >= 1) v = s * dr[5 - m];
      }
      Wy[m * plane + k] = v;
    }
  }
}

// lq[c] = sum_m Wx[m][c - (m-3)] + Wy[m][c - (m-3) rows] (periodic), and
// perpgrad^T of lamU: Ux[j,i] = sx (psi[j,i] - psi[j+1,i]), Uy[j,i] =
// sy (psi[j,i+1] - psi[j,i])  =>  lpsi[j,i] = sx (lUx[j,i] - lUx[j-1,i]) +
// sy (lUy[j,i-1] - lUy[j,i]) (periodic)
template <class T>
__global__ __launch_bounds__(256) void k_adj_gather_perp(AdjArgs a, const T* __restrict__ Wx,
                                                         const T* __restrict__ Wy,
                                                         const T* __restrict__ lUx,
                                                         const T* __restrict__ lUy,
                                                         double sx, double sy,
                                                         T* __restrict__ lq,
                                                         T* __restrict__ lpsi,
                                                         int64_t plane) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)a.ny * a.nx) return;
  const int y = (int)(idx / a.nx), x = (int)(idx - (int64_t)y * a.nx);
  const int64_t P = a.NXP;
  const int64_t k = (int64_t)(y + 3) * P + x + 3;
  T acc = T(0);
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const int xf = wrapi(x - (m - 3), a.nx);
    acc += Wx[m * plane + (int64_t)(y + 3) * P + xf + 3];
  }
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const int yf = wrapi(y - (m - 3), a.ny);
    acc += Wy[m * plane + (int64_t)(yf + 3) * P + x + 3];
  }
  lq[k] = acc;
  const int64_t km = (int64_t)(wrapi(y - 1, a.ny) + 3) * P + x + 3;
  const int64_t kw = (int64_t)(y + 3) * P + wrapi(x - 1, a.nx) + 3;
  lpsi[k] = T(sx) * (lUx[k] - lUx[km]) + T(sy) * (lUy[kw] - lUy[k]);
}

// out = base + lq + c2v^T lb, c2v^T lb [j,i] = 0.25 (lb[j,i] + lb[j+1,i] +
// lb[j,i+1] + lb[j+1,i+1]) (periodic); interior only (out may be base)
template <class T>
__global__ __launch_bounds__(256) void k_adj_combine(AdjArgs a, const T* base,
                                                     const T* __restrict__ lq,
                                                     const T* __restrict__ lb, T* out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)a.ny * a.nx) return;
  const int y = (int)(idx / a.nx), x = (int)(idx - (int64_t)y * a.nx);
  const int64_t P = a.NXP;
  const int yp = wrapi(y + 1, a.ny), xp = wrapi(x + 1, a.nx);
  const int64_t k = (int64_t)(y + 3) * P + x + 3;
  const T s = ((lb[k] + lb[(int64_t)(yp + 3) * P + x + 3]) +
               lb[(int64_t)(y + 3) * P + xp + 3]) +
              lb[(int64_t)(yp + 3) * P + xp + 3];
  out[k] = base[k] + (lq[k] + T(0.25) * s);
}

}  // namespace f2d

using namespace f2d;

extern "C" {

int f2d_adj_stage(int dtype, const f2d_grid* g, const void* q, const void* Ux,
                  const void* Uy, const void* lam0, const void* lam1, const void* lam2,
                  double c0, double c1, double c2, void* lUx, void* lUy, void* work,
                  void* lq, void* lpsi, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(g && g->halowidth == 3 && g->xperiodic && g->yperiodic,
                "adjoint: doubly periodic grid with halo 3 only");
    F2D_REQUIRE(q && Ux && Uy && lam0 && lUx && lUy && work && lq && lpsi,
                "NULL argument");
    F2D_REQUIRE(lam1 || !lam2, "lam2 given without lam1");
    AdjArgs a{(int)g->ny, (int)g->nx, (int)g->nx + 6};
    const int64_t plane = (g->ny + 6) * (g->nx + 6);
    const int64_t n = g->ny * g->nx;
    const double sx = 1.0 / (g->dy * g->dy), sy = 1.0 / (g->dx * g->dx);
    hipStream_t s = S(stream);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      T* W = (T*)work;
      LamComb<T> lc{(const T*)lam0, (const T*)lam1, (const T*)lam2, T(c0), T(c1), T(c2)};
      k_adj_faces<T><<<grid1d(n, 256), 256, 0, s>>>(a, (const T*)q, (const T*)Ux,
                                                    (const T*)Uy, lc, (T*)lUx, (T*)lUy, W,
                                                    W + 6 * plane, plane);
      k_adj_gather_perp<T><<<grid1d(n, 256), 256, 0, s>>>(a, W, W + 6 * plane,
                                                          (const T*)lUx, (const T*)lUy, sx,
                                                          sy, (T*)lq, (T*)lpsi, plane);
    };
    if (dtype == F2D_F64)
      run(double());
    else if (dtype == F2D_F32)
      run(float());
    else
      throw Error{"dtype must be F2D_F64 or F2D_F32"};
  });
}

int f2d_adj_combine(int dtype, const f2d_grid* g, const void* base, const void* lq,
                    const void* lb, void* out, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(g && g->xperiodic && g->yperiodic, "adjoint: doubly periodic only");
    F2D_REQUIRE(base && lq && lb && out, "NULL argument");
    AdjArgs a{(int)g->ny, (int)g->nx, (int)g->nx + 6};
    const int64_t n = g->ny * g->nx;
    if (dtype == F2D_F64)
      k_adj_combine<double><<<grid1d(n, 256), 256, 0, S(stream)>>>(
          a, (const double*)base, (const double*)lq, (const double*)lb, (double*)out);
    else if (dtype == F2D_F32)
      k_adj_combine<float><<<grid1d(n, 256), 256, 0, S(stream)>>>(
          a, (const float*)base, (const float*)lq, (const float*)lb, (float*)out);
    else
      throw Error{"dtype must be F2D_F64 or F2D_F32"};
  });
}

}  // extern "C"
