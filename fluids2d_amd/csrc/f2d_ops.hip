// f2d_ops.hip -- the C-grid point operators of the shallow-water models
// (operators.py), one kernel per numpy statement group, over the padded
// (ny+6) x (nx+6) arrays.
//
// Each kernel writes exactly the entries the reference's slices assign and
// leaves the others untouched (their old values stay, as in numpy), applies
// masks by multiplication (so NaN and -0 propagate as numpy's do) and
// evaluates sums left to right.  In FP32 the Python float constants are
// rounded to float first, as numpy does for a float32 array times a Python
// scalar; products with float64 arrays (slip coefficients, 1/coef) are formed
// in double and rounded once, as numpy's in-place same-kind casting does.
#include "f2d_common.hpp"

namespace f2d {

__device__ __forceinline__ bool rc_of(int64_t k, int NX, int& j, int& i) {
  j = (int)(k / NX);
  i = (int)(k - (int64_t)j * NX);
  return true;
}

// addcoriolis, operators.py:32-39
template <class T>
__global__ void k_coriolis(const T* __restrict__ Ux, const T* __restrict__ Uy,
                           T* __restrict__ dux, T* __restrict__ duy, double f,
                           int NY, int NX) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)NY * NX) return;
  int j, i;
  rc_of(k, NX, j, i);
  const T ff = T(f);
  if (j < NY - 1 && i >= 1 && i < NX - 1)
    dux[k] = dux[k] + ff * (((Uy[k - 1] + Uy[k + NX - 1]) + Uy[k]) + Uy[k + NX]);
  if (j >= 1 && j < NY - 1 && i < NX - 1)
    duy[k] = duy[k] - ff * (((Ux[k - NX] + Ux[k - NX + 1]) + Ux[k]) + Ux[k + 1]);
}

// addgrad (vector case), operators.py:49-56
template <class T>
__global__ void k_addgrad(const T* __restrict__ phi, const int8_t* __restrict__ mx,
                          const int8_t* __restrict__ my, T* __restrict__ dux,
                          T* __restrict__ duy, int NY, int NX) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)NY * NX) return;
  int j, i;
  rc_of(k, NX, j, i);
  if (i >= 1) dux[k] = dux[k] - (phi[k] - phi[k - 1]) * T(mx[k]);
  if (j >= 1) duy[k] = duy[k] - (phi[k] - phi[k - NX]) * T(my[k]);
}

// sharp, operators.py:59-64 (sx = 1/dx**2, sy = 1/dy**2)
template <class T>
__global__ void k_sharp(const T* __restrict__ ux, const T* __restrict__ uy,
                        T* __restrict__ Ux, T* __restrict__ Uy, double sx,
                        double sy, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  Ux[k] = ux[k] * T(sx);
  Uy[k] = uy[k] * T(sy);
}

// compute_vorticity, operators.py:67-77: omega[1:, :] = -diff(u.x, axis=0);
// omega[:, 1:] += diff(u.y, axis=1); omega *= slipcoef (row 0 keeps its
// old value before the += and the product)
template <class T>
__global__ void k_vorticity(const T* __restrict__ ux, const T* __restrict__ uy,
                            const double* __restrict__ slip, T* __restrict__ w,
                            int NY, int NX) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)NY * NX) return;
  int j, i;
  rc_of(k, NX, j, i);
  T v = j >= 1 ? -(ux[k] - ux[k - NX]) : w[k];
  if (i >= 1) v = v + (uy[k] - uy[k - 1]);
  w[k] = T((double)v * slip[k]);
}

// ke *= msk * 0.5, operators.py:99-101
template <class T>
__global__ void k_ke_finish(T* __restrict__ ke, const int8_t* __restrict__ msk,
                            int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  ke[k] = T((double)ke[k] * ((double)msk[k] * 0.5));
}

// flag[0] = 1 where a wet entry (msk NULL: any entry) is not finite; the
// flag is only ever set here (the caller clears it)
template <class T>
__global__ void k_nonfinite(const T* __restrict__ a, const int8_t* __restrict__ msk,
                            int64_t n, int32_t* __restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if ((msk == nullptr || msk[k]) && !isfinite((double)a[k])) flag[0] = 1;
}

// compute_pressure, operators.py:110-111: p = (g/area) * (h + hb)
template <class T>
__global__ void k_pressure(const T* __restrict__ h, T* __restrict__ p, double goa,
                           double hb, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  p[k] = T(goa) * (h[k] + T(hb));
}

// div, operators.py:104-107: delta[:, :-1] = -diff(F.x, axis=1);
// delta[:-1, :] -= diff(F.y, axis=0); delta *= msk
template <class T>
__global__ void k_div(const T* __restrict__ Fx, const T* __restrict__ Fy,
                      const int8_t* __restrict__ msk, T* __restrict__ d, int NY,
                      int NX) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)NY * NX) return;
  int j, i;
  rc_of(k, NX, j, i);
  T v = i < NX - 1 ? -(Fx[k + 1] - Fx[k]) : d[k];
  if (j < NY - 1) v = v - (Fy[k + NX] - Fy[k]);
  d[k] = v * T(msk[k]);
}

// centerstovertices of (c * scale), operators.py:126-133, with addto; then
// vomega *= mskv
template <class T>
__global__ void k_c2v_scaled(const T* __restrict__ c, double scale,
                             T* __restrict__ v, const int8_t* __restrict__ mv,
                             int addto, int NY, int NX) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)NY * NX) return;
  int j, i;
  rc_of(k, NX, j, i);
  T val = v[k];
  if (j >= 1 && i >= 1) {
    const T sc = T(scale);
    const T s = T(0.25) * ((((c[k - NX - 1] * sc) + (c[k - 1] * sc)) + (c[k - NX] * sc)) +
                           (c[k] * sc));
    val = addto ? val + s : s;
  }
  v[k] = val * T(mv[k]);
}

// verticestocenters of (vh * scale), operators.py:136-141: h[:-1, :-1] =
// (1/coef) * (four vertex values), coef the number of wet vertices (1/0 is
// inf and inf * 0 is NaN, as in numpy); then h *= msk
template <class T>
__global__ void k_v2c_scaled(const T* __restrict__ vh, double scale,
                             const int8_t* __restrict__ mv,
                             const int8_t* __restrict__ msk, T* __restrict__ h,
                             int NY, int NX) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)NY * NX) return;
  int j, i;
  rc_of(k, NX, j, i);
  T val = h[k];
  if (j < NY - 1 && i < NX - 1) {
    const int coef = (((int)mv[k] + (int)mv[k + NX]) + (int)mv[k + 1]) + (int)mv[k + NX + 1];
    const T sc = T(scale);
    const T s = (((vh[k] * sc) + (vh[k + NX] * sc)) + (vh[k + 1] * sc)) + (vh[k + NX + 1] * sc);
    val = T(inv_count(coef) * (double)s);
  }
  h[k] = val * T(msk[k]);
}


// Bulk diagnostics (diagnostics.py:45-66): FP64 sums over every point of
// the padded arrays, as numpy's domsum does.  Per block the partial sums of
// ke, omega^2, omega, Uy xv, Ux yu and msk (6 doubles), then one block
// reduces them in a fixed order (deterministic) and forms
//   ke = S_ke / S_m, ens = 0.5 (S_w2 / S_m) / area^2, vort = (S_w / S_m) / area,
//   angular = (S_uyx / S_m) dx - (S_uxy / S_m) dy,
// with xv = (i - 3 + 1/2) dx, yu = (j - 3 + 1/2) dy (mesh.xy, meshes.py:57-68).
// ke NULL: the cell-centred 1/2 u.U of the contravariant U (covariant
// u = U dx^2 in x, U dy^2 in y), a quarter of the four face products.
struct BulkArgs {
  int NY, NX;
  double dx, dy;
};

__device__ __forceinline__ double bsum_block(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

template <class T>
__global__ __launch_bounds__(256) void k_bulk_part(BulkArgs a, const T* __restrict__ ke,
                                                   const T* __restrict__ w,
                                                   const T* __restrict__ Ux,
                                                   const T* __restrict__ Uy,
                                                   const int8_t* __restrict__ msk,
                                                   double* __restrict__ part) {
  __shared__ double sh[256];
  const int64_t n = (int64_t)a.NY * a.NX;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(k / a.NX), i = (int)(k - (int64_t)j * a.NX);
    const double ux = (double)Ux[k], uy = (double)Uy[k], wv = (double)w[k];
    double kv;
    if (ke) {
      kv = (double)ke[k];
    } else {
      const double uxe = i + 1 < a.NX ? (double)Ux[k + 1] : 0.0;
      const double uyn = j + 1 < a.NY ? (double)Uy[k + a.NX] : 0.0;
      kv = 0.25 * (a.dx * a.dx * (ux * ux + uxe * uxe) + a.dy * a.dy * (uy * uy + uyn * uyn)) *
           (double)msk[k];
    }
    s[0] += kv;
    s[1] += wv * wv;
    s[2] += wv;
    s[3] += uy * ((i - 3 + 0.5) * a.dx);
    s[4] += ux * ((j - 3 + 0.5) * a.dy);
    s[5] += (double)msk[k];
  }
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const double v = bsum_block(s[q], sh);
    if (threadIdx.x == 0) part[(int64_t)q * gridDim.x + blockIdx.x] = v;
  }
}

__global__ __launch_bounds__(256) void k_bulk_fin(BulkArgs a, const double* __restrict__ part,
                                                  int nb, double area, double* __restrict__ out) {
  __shared__ double sh[256];
  double S[6];
  for (int q = 0; q < 6; q++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[(int64_t)q * nb + b];
    S[q] = bsum_block(v, sh);
  }
  if (threadIdx.x == 0) {
    const double m = S[5];
    out[0] = S[0] / m;
    out[1] = 0.5 * (S[1] / m) / (area * area);
    out[2] = (S[2] / m) / area;
    out[3] = (S[3] / m) * a.dx - (S[4] / m) * a.dy;
  }
}

}  // namespace f2d

using namespace f2d;

#define F2D_OPS_DISPATCH(dtype, ...)                                        \
  do {                                                                      \
    if ((dtype) == F2D_F64) {                                               \
      using T = double;                                                     \
      __VA_ARGS__;                                                          \
    } else if ((dtype) == F2D_F32) {                                        \
      using T = float;                                                      \
      __VA_ARGS__;                                                          \
    } else {                                                                \
      throw ::f2d::Error{"dtype must be F2D_F64 or F2D_F32"};               \
    }                                                                       \
  } while (0)

static void check_grid(const f2d_grid* g) {
  F2D_REQUIRE(g && g->halowidth == 3 && g->ny > 0 && g->nx > 0, "bad grid");
}

extern "C" {

int f2d_coriolis(int dtype, const f2d_grid* g, const void* Ux, const void* Uy,
                 void* dux, void* duy, double f, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int NY = (int)g->ny + 6, NX = (int)g->nx + 6;
    F2D_OPS_DISPATCH(dtype, k_coriolis<T><<<grid1d((int64_t)NY * NX, 256), 256, 0,
                                            S(stream)>>>(
                                (const T*)Ux, (const T*)Uy, (T*)dux, (T*)duy, f, NY, NX));
  });
}

int f2d_addgrad(int dtype, const f2d_grid* g, const void* phi, const int8_t* mskx,
                const int8_t* msky, void* dux, void* duy, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int NY = (int)g->ny + 6, NX = (int)g->nx + 6;
    F2D_OPS_DISPATCH(dtype, k_addgrad<T><<<grid1d((int64_t)NY * NX, 256), 256, 0,
                                           S(stream)>>>(
                                (const T*)phi, mskx, msky, (T*)dux, (T*)duy, NY, NX));
  });
}

int f2d_sharp(int dtype, const f2d_grid* g, const void* ux, const void* uy,
              void* Ux, void* Uy, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int64_t n = (g->ny + 6) * (g->nx + 6);
    const double sx = 1.0 / (g->dx * g->dx), sy = 1.0 / (g->dy * g->dy);
    F2D_OPS_DISPATCH(dtype, k_sharp<T><<<grid1d(n, 256), 256, 0, S(stream)>>>(
                                (const T*)ux, (const T*)uy, (T*)Ux, (T*)Uy, sx, sy, n));
  });
}

int f2d_vorticity(int dtype, const f2d_grid* g, const void* ux, const void* uy,
                  const double* slipcoef, void* omega, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int NY = (int)g->ny + 6, NX = (int)g->nx + 6;
    F2D_OPS_DISPATCH(dtype, k_vorticity<T><<<grid1d((int64_t)NY * NX, 256), 256, 0,
                                             S(stream)>>>(
                                (const T*)ux, (const T*)uy, slipcoef, (T*)omega, NY, NX));
  });
}

int f2d_ke_finish(int dtype, const f2d_grid* g, void* ke, const int8_t* msk,
                  void* stream) {
  return guarded([&] {
    check_grid(g);
    const int64_t n = (g->ny + 6) * (g->nx + 6);
    F2D_OPS_DISPATCH(dtype, k_ke_finish<T><<<grid1d(n, 256), 256, 0, S(stream)>>>(
                                (T*)ke, msk, n));
  });
}

int f2d_nonfinite(int dtype, const f2d_grid* g, const void* a, const int8_t* msk,
                  int32_t* flag, void* stream) {
  return guarded([&] {
    check_grid(g);
    F2D_REQUIRE(a && flag, "NULL argument");
    const int64_t n = (g->ny + 6) * (g->nx + 6);
    F2D_OPS_DISPATCH(dtype, k_nonfinite<T><<<grid1d(n, 256), 256, 0, S(stream)>>>(
                                (const T*)a, msk, n, flag));
  });
}

int f2d_pressure(int dtype, const f2d_grid* g, const void* h, void* p,
                 double g_over_area, double hb, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int64_t n = (g->ny + 6) * (g->nx + 6);
    F2D_OPS_DISPATCH(dtype, k_pressure<T><<<grid1d(n, 256), 256, 0, S(stream)>>>(
                                (const T*)h, (T*)p, g_over_area, hb, n));
  });
}

int f2d_div(int dtype, const f2d_grid* g, const void* Fx, const void* Fy,
            const int8_t* msk, void* delta, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int NY = (int)g->ny + 6, NX = (int)g->nx + 6;
    F2D_OPS_DISPATCH(dtype, k_div<T><<<grid1d((int64_t)NY * NX, 256), 256, 0,
                                       S(stream)>>>(
                                (const T*)Fx, (const T*)Fy, msk, (T*)delta, NY, NX));
  });
}

int f2d_c2v_scaled(int dtype, const f2d_grid* g, const void* c, double scale,
                   void* v, const int8_t* mskv, int addto, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int NY = (int)g->ny + 6, NX = (int)g->nx + 6;
    F2D_OPS_DISPATCH(dtype, k_c2v_scaled<T><<<grid1d((int64_t)NY * NX, 256), 256, 0,
                                              S(stream)>>>(
                                (const T*)c, scale, (T*)v, mskv, addto, NY, NX));
  });
}

int f2d_v2c_scaled(int dtype, const f2d_grid* g, const void* vh, double scale,
                   const int8_t* mskv, const int8_t* msk, void* h, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int NY = (int)g->ny + 6, NX = (int)g->nx + 6;
    F2D_OPS_DISPATCH(dtype, k_v2c_scaled<T><<<grid1d((int64_t)NY * NX, 256), 256, 0,
                                              S(stream)>>>(
                                (const T*)vh, scale, mskv, msk, (T*)h, NY, NX));
  });
}

int f2d_bulk_diagnostics(int dtype, const f2d_grid* g, const void* ke, const void* omega,
                         const void* Ux, const void* Uy, const int8_t* msk, double* partials,
                         int nparts, double* out, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(g && omega && Ux && Uy && msk && partials && out, "NULL argument");
    F2D_REQUIRE(nparts >= 1, "nparts must be >= 1");
    BulkArgs a{(int)g->ny + 6, (int)g->nx + 6, g->dx, g->dy};
    const int64_t n = (int64_t)a.NY * a.NX;
    int nb = grid1d(n, 256);
    if (nb > nparts) nb = nparts;
    hipStream_t s = S(stream);
    F2D_OPS_DISPATCH(dtype, k_bulk_part<T><<<nb, 256, 0, s>>>(a, (const T*)ke, (const T*)omega,
                                                               (const T*)Ux, (const T*)Uy, msk,
                                                               partials));
    k_bulk_fin<<<1, 256, 0, s>>>(a, partials, nb, g->dx * g->dy, out);
  });
}

}  // extern "C"
