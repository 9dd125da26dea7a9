// f2d_mesh.hip -- mask and stencil-order logic of the mesh (integer work).
//
//   set_masks   meshes.py:77-86
//   set_order   meshes.py:146-186
//   slipcoef    noslip.py:4-36
//   gindex      elliptic.py:198-203
#include <string>

#include "f2d_common.hpp"

namespace f2d {

thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// mskx[j,i] = msk[j,i]*msk[j,i-1] (i>0), msky likewise in j, mskv the
// product of the four cells around the vertex (meshes.py:78-86).
__global__ void k_set_masks(const int8_t* __restrict__ m, int64_t ny,
                            int64_t nx, int8_t* __restrict__ mx,
                            int8_t* __restrict__ my, int8_t* __restrict__ mv) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ny * nx) return;
  int64_t j = k / nx, i = k - j * nx;
  int a = m[k];
  mx[k] = (i > 0) ? (int8_t)(a * m[k - 1]) : 0;
  my[k] = (j > 0) ? (int8_t)(a * m[k - nx]) : 0;
  mv[k] = (i > 0 && j > 0)
              ? (int8_t)(m[k - nx] * m[k - nx - 1] * a * m[k - 1])
              : 0;
}

// set_order (meshes.py:146-186): order 6 if the 6-cell stencil around the
// face is wet, else 4 if the 4-cell one is, else 2 if the 2-cell one is,
// else 0, capped by maxorder; bounds are checked on the flat index.
__global__ void k_set_order(const int8_t* __restrict__ m, int64_t n,
                            int64_t shift, int maxorder,
                            int8_t* __restrict__ o) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s2, s4, s6;
  if (shift > 0) {
    s2 = (i - shift >= 0) ? m[i - shift] + m[i] : 0;
    s4 = ((i - shift * 2 >= 0) && (i + shift < n))
             ? m[i - shift * 2] + m[i + shift] + s2
             : 0;
    s6 = ((i - shift * 3 >= 0) && (i + 2 * shift < n))
             ? m[i - shift * 3] + m[i + shift * 2] + s4
             : 0;
  } else {
    s2 = (i - shift < n) ? m[i - shift] + m[i] : 0;
    s4 = ((i - shift * 2 < n) && (i + shift >= 0))
             ? m[i - shift * 2] + m[i + shift] + s2
             : 0;
    s6 = ((i - shift * 3 < n) && (i + 2 * shift >= 0))
             ? m[i - shift * 3] + m[i + shift * 2] + s4
             : 0;
  }
  int v = (s6 == 6) ? 6 : ((s4 == 4) ? 4 : ((s2 == 2) ? 2 : 0));
  o[i] = (int8_t)(v < maxorder ? v : maxorder);
}

// get_slipcoef (noslip.py:4-36): coef counts the wet cells around the
// vertex; free slip keeps vertices with 4, no slip those with >= 1.
__global__ void k_slipcoef(const int8_t* __restrict__ m, int64_t ny,
                           int64_t nx, int mode, int sides, int nh,
                           double* __restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ny * nx) return;
  int64_t j = k / nx, i = k - j * nx;
  double c = (double)m[k];
  if (i > 0) c += m[k - 1];
  if (j > 0) c += m[k - nx];
  if (i > 0 && j > 0) c += m[k - nx - 1];
  double freeslip = (c == 4.0) ? 1.0 : 0.0;
  double noslip = c < 1.0 ? c : 1.0;
  double v;
  if (mode == 0) {
    v = freeslip;
  } else if (mode == 1) {
    v = noslip;
  } else {
    bool ns = ((sides & 1) && i == nh) || ((sides & 2) && i == nx - nh) ||
              ((sides & 4) && j == nh) || ((sides & 8) && j == ny - nh);
    v = ns ? noslip : freeslip;
  }
  out[k] = v;
}

// Row-major running index of wet points: one workgroup scans the array in
// chunks of 1024 (set-up only, deterministic).
__global__ __launch_bounds__(1024) void k_gindex(const int8_t* __restrict__ m,
                                                 int64_t n,
                                                 int32_t* __restrict__ G,
                                                 int64_t* __restrict__ count) {
  __shared__ int sh[1024];
  __shared__ int64_t base;
  const int t = threadIdx.x;
  if (t == 0) base = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < n; c0 += 1024) {
    int64_t i = c0 + t;
    int v = (i < n && m[i]) ? 1 : 0;
    sh[t] = v;
    __syncthreads();
    for (int s = 1; s < 1024; s <<= 1) {  // inclusive Hillis-Steele scan
      int a = (t >= s) ? sh[t - s] : 0;
      __syncthreads();
      sh[t] += a;
      __syncthreads();
    }
    if (i < n) G[i] = v ? (int32_t)(base + sh[t] - 1) : -1;
    __syncthreads();
    if (t == 0) base += sh[1023];
    __syncthreads();
  }
  if (t == 0) *count = base;
}

}  // namespace f2d

using namespace f2d;

extern "C" {

int f2d_abi_version(void) { return 8; }

const char* f2d_last_error(void) { return g_last_error.c_str(); }

int f2d_device_synchronize(void) {
  return guarded([&] { F2D_HIP(hipDeviceSynchronize()); });
}

int f2d_set_masks(const int8_t* msk, int64_t ny, int64_t nx, int8_t* mskx,
                  int8_t* msky, int8_t* mskv, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(ny > 0 && nx > 0, "f2d_set_masks: empty array");
    k_set_masks<<<grid1d(ny * nx, 256), 256, 0, S(stream)>>>(msk, ny, nx, mskx,
                                                              msky, mskv);
  });
}

int f2d_set_order(const int8_t* m, int64_t n, int64_t shift, int maxorder,
                  int8_t* o, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(n > 0, "f2d_set_order: empty array");
    F2D_REQUIRE(shift != 0, "f2d_set_order: shift must be non-zero");
    k_set_order<<<grid1d(n, 256), 256, 0, S(stream)>>>(m, n, shift, maxorder,
                                                        o);
  });
}

int f2d_slipcoef(const int8_t* msk, int64_t ny, int64_t nx, int mode,
                 int sides, int halowidth, double* coef, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(mode >= 0 && mode <= 2, "f2d_slipcoef: bad mode");
    k_slipcoef<<<grid1d(ny * nx, 256), 256, 0, S(stream)>>>(
        msk, ny, nx, mode, sides, halowidth, coef);
  });
}

int f2d_gindex(const int8_t* m, int64_t n, int32_t* G, int64_t* count_host,
               void* stream) {
  return guarded([&] {
    F2D_REQUIRE(n > 0, "f2d_gindex: empty array");
    int64_t* dcount = nullptr;
    F2D_HIP(hipMallocAsync((void**)&dcount, sizeof(int64_t), S(stream)));
    k_gindex<<<1, 1024, 0, S(stream)>>>(m, n, G, dcount);
    int64_t c = 0;
    F2D_HIP(hipMemcpyAsync(&c, dcount, sizeof(int64_t), hipMemcpyDeviceToHost,
                           S(stream)));
    F2D_HIP(hipFreeAsync(dcount, S(stream)));
    F2D_HIP(hipStreamSynchronize(S(stream)));
    if (count_host) *count_host = c;
  });
}

}  // extern "C"
