// f2d_common.hpp -- shared device/host helpers of the fluids2d_amd library.
//
// All kernels are compiled with -ffp-contract=off: the FP64 path reproduces
// the reference's (numba, no fastmath) IEEE double arithmetic operation by
// operation, so FMA contraction must not merge a multiply and an add.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fluids2d_amd.h"

namespace f2d {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

struct Error {
  std::string msg;
};

#define F2D_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw ::f2d::Error{std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define F2D_REQUIRE(cond, msg)                                  \
  do {                                                          \
    if (!(cond)) throw ::f2d::Error{std::string(msg)};          \
  } while (0)

// ---------------------------------------------------------------------------
// member batching: kernel launches recorded instead of run
// ---------------------------------------------------------------------------
// Between f2d_record_begin and f2d_record_end (f2d_batch.hip) the entry
// points that support it append their kernel launches -- kernel, grid,
// block, dynamic LDS and the by-value argument struct -- to this thread's
// recording instead of launching them; any other entry point refuses to
// run.  The recordings of several members, kernel for kernel alike, become
// one launch per kernel for all of them (blockIdx.y = member), each
// workgroup taking its member's argument struct from a device table.
struct LaunchRec {
  const void* fn;  // the kernel's batched instantiation
  dim3 grid, block;
  unsigned shmem;
  std::vector<unsigned char> args;
};
struct Recording {
  std::vector<LaunchRec> launches;
};
Recording*& recording();  // this thread's recording (NULL: launches run)
bool& recordable();       // set while an entry point that can be recorded runs

// Launch kernel `plain`(a, NULL), or record `batched` with a.  Both are
// instantiations of one kernel template whose batched form replaces `a` by
// table[blockIdx.y] (its second argument).
template <class A>
inline void launch_rec(void (*plain)(A, const A*), void (*batched)(A, const A*), dim3 grid,
                       dim3 block, unsigned shmem, hipStream_t s, const A& a) {
  if (Recording* r = recording()) {
    LaunchRec l{(const void*)batched, grid, block, shmem, std::vector<unsigned char>(sizeof(A))};
    std::memcpy(l.args.data(), &a, sizeof(A));
    r->launches.push_back(std::move(l));
    return;
  }
  hipLaunchKernelGGL(plain, grid, block, shmem, s, a, (const A*)nullptr);
}

// Wrap an entry point body: exceptions become a status code + message.
template <class F>
int guarded(F&& f) {
  try {
    if (recording() && !recordable())
      throw Error{"this entry point cannot be recorded into a member batch"};
    f();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(std::string("kernel launch: ") + hipGetErrorString(e));
      return -2;
    }
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return -1;
  } catch (const std::exception& e) {
    set_error(e.what());
    return -3;
  }
}

// guarded() for an entry point whose launches all go through launch_rec
template <class F>
int guarded_rec(F&& f) {
  recordable() = true;
  const int rc = guarded(f);
  recordable() = false;
  return rc;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int grid1d(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  return (int)(b < 1 ? 1 : b);
}

// ---------------------------------------------------------------------------
// pointwise reconstructions (weno.py:21-43, 75-134, 167-197, 255-300),
// evaluation order as in the Python source; T = double reproduces it exactly.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T ce2(T qm, T qp) {
  return (qm + qp) * T(0.5);
}
template <class T>
__device__ __forceinline__ T ce4(T qmm, T qm, T qp, T qpp) {
  return ((-qmm + T(7) * (qm + qp)) - qpp) / T(12);
}
template <class T>
__device__ __forceinline__ T ce6(T qmmm, T qmm, T qm, T qp, T qpp, T qppp) {
  return (((qmmm - T(8) * (qmm + qpp)) + T(37) * (qm + qp)) + qppp) / T(60);
}
template <class T>
__device__ __forceinline__ T up3(T qm, T q0, T qp) {
  return ((T(5) * q0 + T(2) * qp) - qm) / T(6);
}
template <class T>
__device__ __forceinline__ T up5(T qmmm, T qmm, T qm, T qp, T qpp) {
  return ((((T(2) * qmmm - T(13) * qmm) + T(47) * qm) + T(27) * qp) -
          T(3) * qpp) /
         T(60);
}

// weno.py:75-100
template <class T>
__device__ __forceinline__ T weno3z(T qm, T q0, T qp) {
  const T eps = T(1e-14);
  T qi1 = T(-1. / 2.) * qm + T(3. / 2.) * q0;
  T qi2 = T(1. / 2.) * (q0 + qp);
  T d1 = q0 - qm, d2 = qp - q0;
  T beta1 = d1 * d1, beta2 = d2 * d2;
  T tau = fabs(beta2 - beta1);
  T w1 = T(1. / 3.) * (T(1) + tau / (beta1 + eps));
  T w2 = T(2. / 3.) * (T(1) + tau / (beta2 + eps));
  return (w1 * qi1 + w2 * qi2) / (w1 + w2);
}

// weno.py:103-134
template <class T>
__device__ __forceinline__ T cweno3z(T U, T qmm, T qm, T qp, T qpp) {
  const T eps = T(1e-14);
  T qi1 = T(-1. / 2.) * qmm + T(3. / 2.) * qm;
  T qi2 = T(1. / 2.) * (qm + qp);
  T qi3 = T(-1. / 2.) * qpp + T(3. / 2.) * qp;
  T a = (U > T(0)) ? (qm - qmm) : (qp - qpp);
  T b = qp - qm;
  T beta1 = a * a, beta2 = b * b;
  T tau = fabs(beta2 - beta1);
  T w1 = T(1. / 3.) * (T(1) + tau / (beta1 + eps));
  T w2 = T(2. / 3.) * (T(1) + tau / (beta2 + eps));
  return ((w1 * (qi1 + qi3)) * T(0.5) + w2 * qi2) / (w1 + w2);
}

// weno.py:167-197
template <class T>
__device__ __forceinline__ T weno5z(T qmm, T qm, T q0, T qp, T qpp) {
  const T eps = T(1e-16);
  T qi1 = (T(1. / 3.) * qmm - T(7. / 6.) * qm) + T(11. / 6.) * q0;
  T qi2 = (T(-1. / 6.) * qm + T(5. / 6.) * q0) + T(1. / 3.) * qp;
  T qi3 = (T(1. / 3.) * q0 + T(5. / 6.) * qp) - T(1. / 6.) * qpp;
  const T k1 = T(13. / 12.), k2 = T(0.25);
  T a1 = (qmm - T(2) * qm) + q0, b1 = (qmm - T(4) * qm) + T(3) * q0;
  T a2 = (qm - T(2) * q0) + qp, b2 = qm - qp;
  T a3 = (q0 - T(2) * qp) + qpp, b3 = (T(3) * q0 - T(4) * qp) + qpp;
  T beta1 = k1 * (a1 * a1) + k2 * (b1 * b1);
  T beta2 = k1 * (a2 * a2) + k2 * (b2 * b2);
  T beta3 = k1 * (a3 * a3) + k2 * (b3 * b3);
  T tau5 = fabs(beta1 - beta3);
  T w1 = T(0.1) * (T(1) + tau5 / (beta1 + eps));
  T w2 = T(0.6) * (T(1) + tau5 / (beta2 + eps));
  T w3 = T(0.3) * (T(1) + tau5 / (beta3 + eps));
  return ((w1 * qi1 + w2 * qi2) + w3 * qi3) / ((w1 + w2) + w3);
}

// FP32 forms of the reconstructions above.  The FP32 path has no bit-exact
// target (its bar is the FP64 result within 1e-5 after 1000 steps), so
// these use fused multiply-adds and the hardware reciprocal (v_rcp_f32,
// 1 ulp) instead of IEEE divisions: about half the VALU instructions of a
// face.  The divisors are normal numbers: beta + eps >= eps and the weight
// sum >= 1.  The FP64 templates above are untouched.
__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

template <>
__device__ __forceinline__ float weno5z<float>(float qmm, float qm, float q0, float qp,
                                               float qpp) {
  const float eps = 1e-16f;
  const float qi1 = fmaf(11.f / 6.f, q0, fmaf(1.f / 3.f, qmm, (-7.f / 6.f) * qm));
  const float qi2 = fmaf(1.f / 3.f, qp, fmaf(-1.f / 6.f, qm, (5.f / 6.f) * q0));
  const float qi3 = fmaf(-1.f / 6.f, qpp, fmaf(1.f / 3.f, q0, (5.f / 6.f) * qp));
  const float k1 = 13.f / 12.f, k2 = 0.25f;
  const float a1 = fmaf(-2.f, qm, qmm) + q0, b1 = fmaf(3.f, q0, fmaf(-4.f, qm, qmm));
  const float a2 = fmaf(-2.f, q0, qm) + qp, b2 = qm - qp;
  const float a3 = fmaf(-2.f, qp, q0) + qpp, b3 = fmaf(3.f, q0, fmaf(-4.f, qp, qpp));
  const float beta1 = fmaf(k1, a1 * a1, k2 * (b1 * b1));
  const float beta2 = fmaf(k1, a2 * a2, k2 * (b2 * b2));
  const float beta3 = fmaf(k1, a3 * a3, k2 * (b3 * b3));
  const float tau5 = fabsf(beta1 - beta3);
  const float w1 = fmaf(0.1f * tau5, rcp32(beta1 + eps), 0.1f);
  const float w2 = fmaf(0.6f * tau5, rcp32(beta2 + eps), 0.6f);
  const float w3 = fmaf(0.3f * tau5, rcp32(beta3 + eps), 0.3f);
  const float num = fmaf(w3, qi3, fmaf(w2, qi2, w1 * qi1));
  return num * rcp32((w1 + w2) + w3);
}

template <>
__device__ __forceinline__ float weno3z<float>(float qm, float q0, float qp) {
  const float eps = 1e-14f;
  const float qi1 = fmaf(-0.5f, qm, 1.5f * q0);
  const float qi2 = 0.5f * (q0 + qp);
  const float d1 = q0 - qm, d2 = qp - q0;
  const float beta1 = d1 * d1, beta2 = d2 * d2;
  const float tau = fabsf(beta2 - beta1);
  const float w1 = fmaf((1.f / 3.f) * tau, rcp32(beta1 + eps), 1.f / 3.f);
  const float w2 = fmaf((2.f / 3.f) * tau, rcp32(beta2 + eps), 2.f / 3.f);
  return fmaf(w2, qi2, w1 * qi1) * rcp32(w1 + w2);
}

template <>
__device__ __forceinline__ float up3<float>(float qm, float q0, float qp) {
  return fmaf(5.f / 6.f, q0, fmaf(1.f / 3.f, qp, (-1.f / 6.f) * qm));
}
template <>
__device__ __forceinline__ float up5<float>(float qmmm, float qmm, float qm, float qp,
                                            float qpp) {
  return fmaf(-3.f / 60.f, qpp,
              fmaf(27.f / 60.f, qp,
                   fmaf(47.f / 60.f, qm, fmaf(-13.f / 60.f, qmm, (2.f / 60.f) * qmmm))));
}
template <>
__device__ __forceinline__ float ce4<float>(float qmm, float qm, float qp, float qpp) {
  return fmaf(7.f / 12.f, qm + qp, (-1.f / 12.f) * (qmm + qpp));
}
template <>
__device__ __forceinline__ float ce6<float>(float qmmm, float qmm, float qm, float qp,
                                            float qpp, float qppp) {
  return fmaf(37.f / 60.f, qm + qp,
              fmaf(-8.f / 60.f, qmm + qpp, (1.f / 60.f) * (qmmm + qppp)));
}

// weno.py:255-300 (cweno5z_v0)
template <class T>
__device__ __forceinline__ T cweno5z(T qmmm, T qmm, T qm, T qp, T qpp,
                                     T qppp) {
  const T eps = T(1e-16);
  T qi1 = (T(1. / 3.) * qmmm - T(7. / 6.) * qmm) + T(11. / 6.) * qm;
  T qi2 = (T(-1. / 6.) * qmm + T(5. / 6.) * qm) + T(1. / 3.) * qp;
  T qi3 = (T(1. / 3.) * qm + T(5. / 6.) * qp) - T(1. / 6.) * qpp;
  T qi4 = (T(1. / 3.) * qp + T(5. / 6.) * qm) - T(1. / 6.) * qmm;
  T qi5 = (T(-1. / 6.) * qpp + T(5. / 6.) * qp) + T(1. / 3.) * qm;
  T qi6 = (T(1. / 3.) * qppp - T(7. / 6.) * qpp) + T(11. / 6.) * qp;
  const T k1 = T(13. / 12.), k2 = T(0.25);
  T a1 = (qmmm - T(2) * qmm) + qm, b1 = (qmmm - T(4) * qmm) + T(3) * qm;
  T a2 = (qmm - T(2) * qm) + qp, b2 = qmm - qp;
  T a3 = (qm - T(2) * qp) + qpp, b3 = (T(3) * qm - T(4) * qp) + qpp;
  T beta1 = k1 * (a1 * a1) + k2 * (b1 * b1);
  T beta2 = k1 * (a2 * a2) + k2 * (b2 * b2);
  T beta3 = k1 * (a3 * a3) + k2 * (b3 * b3);
  T tau5 = fabs(beta1 - beta3);
  T w1 = T(0.1) * (T(1) + tau5 / (beta1 + eps));
  T w2 = T(0.6) * (T(1) + tau5 / (beta2 + eps));
  T w3 = T(0.3) * (T(1) + tau5 / (beta3 + eps));
  return ((w1 * (qi1 + qi6) + w2 * (qi2 + qi5)) + w3 * (qi3 + qi4)) /
         (T(2) * ((w1 + w2) + w3));
}

// Upwind selection, weno.py:303-343.  The WENO branch selects the stencil
// operands first and evaluates weno5z once, so a wavefront whose lanes see
// both signs of U does not run both evaluations.
template <class T>
__device__ __forceinline__ T f1(int method, T U, T qm, T qp) {
  if (method >= F2D_CENTERED) return ce2(qm, qp);
  return U > T(0) ? qm : qp;
}

template <class T>
__device__ __forceinline__ T f3(int method, T U, T qmm, T qm, T qp, T qpp) {
  const bool pos = U > T(0);
  switch (method) {
    case F2D_WENO:
      return pos ? weno3z(qmm, qm, qp) : weno3z(qpp, qp, qm);
    case F2D_UPWIND:
      return pos ? up3(qmm, qm, qp) : up3(qpp, qp, qm);
    case F2D_CENTERED:
      return ce4(qmm, qm, qp, qpp);
    default:
      return cweno3z(U, qmm, qm, qp, qpp);
  }
}

template <class T>
__device__ __forceinline__ T f5(int method, T U, T qmmm, T qmm, T qm, T qp,
                                T qpp, T qppp) {
  const bool pos = U > T(0);
  switch (method) {
    case F2D_WENO: {
      T a = pos ? qmmm : qppp, b = pos ? qmm : qpp, c = pos ? qm : qp;
      T d = pos ? qp : qm, e = pos ? qpp : qmm;
      return weno5z(a, b, c, d, e);
    }
    case F2D_UPWIND:
      return pos ? up5(qmmm, qmm, qm, qp, qpp) : up5(qppp, qpp, qp, qm, qmm);
    case F2D_CENTERED:
      return ce6(qmmm, qmm, qm, qp, qpp, qppp);
    default:
      return cweno5z(qmmm, qmm, qm, qp, qpp, qppp);
  }
}

// Specialised WENO selection used by the fused kernels (method known at
// compile time through the template argument).
template <class T, int M>
__device__ __forceinline__ T flux_by_order(int o, T U, T q_3, T q_2, T q_1,
                                           T q0, T q1, T q2) {
  // stencil around the face between q_1 and q0 (compflux, weno.py:352-363)
  if (o > 4) return f5<T>(M, U, q_3, q_2, q_1, q0, q1, q2);
  if (o > 2) return f3<T>(M, U, q_2, q_1, q0, q1);
  if (o > 0) return f1<T>(M, U, q_1, q0);
  return T(0);
}

// ---------------------------------------------------------------------------
// RK stage coefficients evaluated from dt exactly as the reference's Python
// expressions (integrators.py:82-124).  Returns the number of coefficients.
// ---------------------------------------------------------------------------
// 1.0 / c for the number c of wet vertices about a cell (verticestocenters,
// operators.py:136-141), bit for bit the IEEE division: the quotients of
// 1 .. 4 are the correctly rounded constants, 1 / 0 = +inf (numpy's
// (1/0) * 0 = NaN follows), other counts divide
__device__ __forceinline__ double inv_count(int c) {
  if (c >= 1 && c <= 4)
    return c == 4 ? 0.25 : (c == 2 ? 0.5 : (c == 1 ? 1.0 : 1.0 / 3.0));
  return c == 0 ? __builtin_inf() : 1.0 / (double)c;
}

__host__ __device__ inline int rk_coefs(int integrator, int stage, double dt,
                                        double c[3]) {
  c[0] = c[1] = c[2] = 0.0;
  if (integrator == F2D_RK3) {
    if (stage == 0) { c[0] = dt; return 1; }
    if (stage == 1) { c[0] = (-3.0 * dt) / 4.0; c[1] = dt / 4.0; return 2; }
    c[0] = (-dt) / 12.0; c[1] = (-dt) / 12.0; c[2] = (2.0 * dt) / 3.0;
    return 3;
  }
  if (integrator == F2D_EF) { c[0] = dt; return 1; }
  // F2D_ENRK3
  if (stage == 0) { c[0] = dt / 3.0; return 1; }
  if (stage == 1) {
    c[0] = ((-dt) / 3.0) - ((5.0 * dt) / 48.0);
    c[1] = (15.0 * dt) / 16.0;
    return 2;
  }
  c[0] = ((5.0 * dt) / 48.0) + (dt / 10.0);
  c[1] = ((-7.0) * dt) / 16.0;
  c[2] = (2.0 * dt) / 5.0;
  return 3;
}

// y + ((0 + c0 x0) + c1 x1) + c2 x2 with Python's left-to-right sum
// (integrators.py:174); the leading 0 + v is v.
template <class T>
__device__ __forceinline__ T rk_combine(int nc, const double c[3], T x0, T x1,
                                        T x2) {
  T s = T(c[0]) * x0;
  if (nc > 1) s = s + T(c[1]) * x1;
  if (nc > 2) s = s + T(c[2]) * x2;
  return s;
}

// ---------------------------------------------------------------------------
// Block reduction of the grid reductions (every block writes its partial;
// a following kernel or host step reduces the partials in a fixed order).
// ---------------------------------------------------------------------------
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

}  // namespace f2d
