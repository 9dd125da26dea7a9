// f2d_stencil.hip -- WENO flux loops, the fused transport stage and the
// C-grid point operators of the eulerpsi/advection/qg paths.
//
//   flat loops          weno.py:346-405, 412-436
//   fused stage         operators.py:21-29,104-107 + meshes.py:106-111 +
//                       integrators.py:92-124,154-196
//   c2v / perpgrad      operators.py:126-133, 144-149
//   halo fill           meshes.py:135-143 (+ y-periodic extension)
//   CFL time step       model.py:71-87
#include <map>
#include <mutex>

#include "f2d_common.hpp"

namespace f2d {

// numba wraps negative indices; an index past the end (never reached with
// orders from set_order) is wrapped too so that no access leaves the array.
__device__ __forceinline__ int64_t wr(int64_t i, int64_t n) {
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

// ---------------------------------------------------------------------------
// flat-index loops, one thread per point
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_compflux(T* __restrict__ flx, const T* __restrict__ U,
                           const T* __restrict__ q, const int8_t* __restrict__ o,
                           int64_t n, int64_t s, int method) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int oi = o[i];
  T u = U[i], r;
  if (oi > 4)
    r = f5<T>(method, u, q[wr(i - 3 * s, n)], q[wr(i - 2 * s, n)],
              q[wr(i - s, n)], q[i], q[wr(i + s, n)], q[wr(i + 2 * s, n)]) *
        u;
  else if (oi > 2)
    r = f3<T>(method, u, q[wr(i - 2 * s, n)], q[wr(i - s, n)], q[i],
              q[wr(i + s, n)]) *
        u;
  else if (oi > 0)
    r = f1<T>(method, u, q[wr(i - s, n)], q[i]) * u;
  else
    r = T(0);
  flx[i] = r;
}

template <class T>
__global__ void k_vortexforce(T* __restrict__ du, const T* __restrict__ V,
                              const T* __restrict__ q,
                              const int8_t* __restrict__ o, int64_t n,
                              int64_t s, int64_t s2, int sign, int method) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int oi = o[i];
  if (oi <= 0) {
    du[i] = T(0);
    return;
  }
  T Vm = T(0.25) * (((V[i] + V[wr(i + s, n)]) + V[wr(i - s2, n)]) +
                    V[wr(i + s - s2, n)]);
  T f;
  if (oi > 4)
    f = f5<T>(method, Vm, q[wr(i - 2 * s, n)], q[wr(i - s, n)], q[i],
              q[wr(i + s, n)], q[wr(i + 2 * s, n)], q[wr(i + 3 * s, n)]);
  else if (oi > 2)
    f = f3<T>(method, Vm, q[wr(i - s, n)], q[i], q[wr(i + s, n)],
              q[wr(i + 2 * s, n)]);
  else
    f = f1<T>(method, Vm, q[i], q[wr(i + s, n)]);
  du[i] = (T(sign) * f) * Vm;
}

template <class T>
__global__ void k_innerproduct(T* __restrict__ ke, const T* __restrict__ U,
                               const T* __restrict__ q,
                               const int8_t* __restrict__ o, int64_t n,
                               int64_t s, int method) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int oi = o[i];
  if (oi <= 0) return;
  T Um = T(0.5) * (U[i] + U[wr(i + s, n)]);
  T f;
  if (oi > 4)
    f = f5<T>(method, Um, q[wr(i - 2 * s, n)], q[wr(i - s, n)], q[i],
              q[wr(i + s, n)], q[wr(i + 2 * s, n)], q[wr(i + 3 * s, n)]);
  else if (oi > 2)
    f = f3<T>(method, Um, q[wr(i - s, n)], q[i], q[wr(i + s, n)],
              q[wr(i + 2 * s, n)]);
  else
    f = f1<T>(method, Um, q[i], q[wr(i + s, n)]);
  ke[i] = ke[i] + f * Um;
}

__device__ __forceinline__ double nanmax(double a, double b) {
  return (a > b || a != a) ? a : b;
}

// block max of v over the 256 threads (NaN-propagating, order-free):
// butterflies within each 64-lane wave, then the 4 wave results
__device__ __forceinline__ double block_nanmax256(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = nanmax(nanmax(sh[0], sh[1]), nanmax(sh[2], sh[3]));
  __syncthreads();
  return r;
}

// The CFL step of model.py:84-87 from k_perpgrad's block maxima part[0..nb)
// of |Ux| and part[nb..2nb) of |Uy|: dt = min(cfl / (max|Ux| + max|Uy| +
// 1e-99), dtmax).  The maxima are exact and order-free, so every caller
// (k_dt_fin's one workgroup, or each transport workgroup) gets the same dt.
// Called by all 256 threads of the workgroup.
// this thread's share (entries t, t + 256, ...) of the maxima of the block
// maxima part[0..nb) and part[nb..2nb); its loads are all issued before
// any is used
__device__ __forceinline__ void cfl_partial_max(const double* __restrict__ part, int nb,
                                                double& ax, double& ay) {
  constexpr int U = 8;
  for (int b0 = 0; b0 < nb; b0 += 256 * U) {
    double vx[U], vy[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int b = b0 + u * 256 + (int)threadIdx.x;
      vx[u] = b < nb ? part[b] : 0.0;
      vy[u] = b < nb ? part[nb + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      ax = nanmax(ax, vx[u]);
      ay = nanmax(ay, vy[u]);
    }
  }
}

__device__ __forceinline__ double cfl_of(double ax, double ay, double cfl, double dtmax) {
  const double v = cfl / ((ax + ay) + 1e-99);
  return (dtmax < v) ? dtmax : v;  // Python min(v, dtmax)
}

__device__ __forceinline__ double cfl_step256(const double* __restrict__ part, int nb,
                                              double cfl, double dtmax, double* sh) {
  double ax = 0.0, ay = 0.0;
  cfl_partial_max(part, nb, ax, ay);
  ax = block_nanmax256(ax, sh);
  ay = block_nanmax256(ay, sh);
  return cfl_of(ax, ay, cfl, dtmax);
}

// ---------------------------------------------------------------------------
// Fused transport stage.
//
// Interior cells [3, 3+ny) x [3, 3+nx) are covered by TY x TX tiles: the tile
// of q with its 3-cell halo is staged in LDS once, every x- and y-face flux
// of the tile is computed once into LDS, then each cell forms
//   ds = ((-(Fx[i+1]-Fx[i])) - (Fy[j+1]-Fy[j])) * msk          (div)
// and q_out = q_in + S(ds_prev, ds)                           (addto).
// Cells of the halo ring are handled by extra blocks: a periodic halo cell
// takes the tendency of its wrapped interior source (what fill() copies), a
// closed-direction halo cell (dry: msk = 0) gets ds = 0.
// ---------------------------------------------------------------------------
template <class T>
struct TransportArgs {
  const T* q;
  T* qo;
  const T* Ux;
  const T* Uy;
  const int8_t* ocx;
  const int8_t* ocy;
  const int8_t* msk;
  const T* ds0;
  const T* ds1;
  T* dso;
  const double* dt;
  int integ, stage;
  int ny, nx, P;  // interior sizes, padded row length
  int xper, yper;
  int tiles_x, ntiles;
  int nhb;  // halo-ring blocks (block indices 0 .. nhb-1)
  double* dt_log;  // nullable: *dt copied there (the step's time step)
  int64_t nhalo;
  // nullable: the stage's dt is the CFL step of these k_perpgrad block
  // maxima (2 cnb entries), formed by every workgroup; workgroup 0 writes
  // it to dt_w (and dt_log) for the later stages
  const double* cpart;
  int cnb;
  double ccfl, cdtmax;
  double* dt_w;
  double dtv;  // the stage's dt, set by the kernel at its start
  // U = perpgrad(psi) formed on the fly instead of read (psi non-NULL):
  // masks, contravariant factors and the periodic halo fill of k_perpgrad
  const T* psi;
  const int8_t* pmx;
  const int8_t* pmy;
  double psx, psy;
  int pfill;
  // U = sharp(u) formed on the fly (uscale != 0): Ux/Uy hold u.x/u.y and
  // the velocity is u.x * usx, u.y * usy, as k_sharp stores it
  int uscale;
  double usx, usy;
};

// U at x face (j, i): read, or perpgrad(psi) with k_perpgrad's arithmetic
// (wrapped source when filling, (-(psi[j+1] - psi[j])) mask, then the
// factor).  Faces read by a stage lie on rows <= ny + 3 and columns
// <= nx + 3, never on the last padded row / column that k_perpgrad leaves
// unchanged.
template <class T>
__device__ __forceinline__ T ux_at(const TransportArgs<T>& a, int j, int i) {
  if (!a.psi) {
    const T u = a.Ux[(int64_t)j * a.P + i];
    return a.uscale ? u * T(a.usx) : u;
  }
  int js = j, is = i;
  if (a.pfill) {
    if (a.xper) is = (i < 3) ? i + a.nx : (i >= a.nx + 3 ? i - a.nx : i);
    if (a.yper) js = (j < 3) ? j + a.ny : (j >= a.ny + 3 ? j - a.ny : j);
  }
  const int64_t ks = (int64_t)js * a.P + is;
  const T u = (-(a.psi[ks + a.P] - a.psi[ks])) * T(a.pmx[ks]);
  return u * T(a.psx);
}

template <class T>
__device__ __forceinline__ T uy_at(const TransportArgs<T>& a, int j, int i) {
  if (!a.psi) {
    const T u = a.Uy[(int64_t)j * a.P + i];
    return a.uscale ? u * T(a.usy) : u;
  }
  int js = j, is = i;
  if (a.pfill) {
    if (a.xper) is = (i < 3) ? i + a.nx : (i >= a.nx + 3 ? i - a.nx : i);
    if (a.yper) js = (j < 3) ? j + a.ny : (j >= a.ny + 3 ? j - a.ny : j);
  }
  const int64_t ks = (int64_t)js * a.P + is;
  const T u = (a.psi[ks + 1] - a.psi[ks]) * T(a.pmy[ks]);
  return u * T(a.psy);
}

template <class T, int M>
__device__ __forceinline__ T xface_global(const TransportArgs<T>& a, int j,
                                          int i) {
  const int64_t k = (int64_t)j * a.P + i;
  const int o = a.ocx[k];
  if (o <= 0) return T(0);
  const T u = ux_at(a, j, i);
  const T* q = a.q + k;
  return flux_by_order<T, M>(o, u, q[-3], q[-2], q[-1], q[0], q[1], q[2]) * u;
}

template <class T, int M>
__device__ __forceinline__ T yface_global(const TransportArgs<T>& a, int j,
                                          int i) {
  const int64_t k = (int64_t)j * a.P + i;
  const int o = a.ocy[k];
  if (o <= 0) return T(0);
  const T u = uy_at(a, j, i);
  const T* q = a.q + k;
  const int64_t P = a.P;
  return flux_by_order<T, M>(o, u, q[-3 * P], q[-2 * P], q[-P], q[0], q[P],
                             q[2 * P]) *
         u;
}

template <class T>
__device__ __forceinline__ void stage_update(const TransportArgs<T>& a,
                                             int64_t k, T ds, T qin) {
  if (a.dso) a.dso[k] = ds;
  if (a.integ < 0) return;
  double c[3];
  const int nc = rk_coefs(a.integ, a.stage, a.dtv, c);
  T x0 = (nc == 1) ? ds : a.ds0[k];
  T x1 = (nc == 2) ? ds : (nc > 2 ? a.ds1[k] : T(0));
  T x2 = (nc == 3) ? ds : T(0);
  a.qo[k] = qin + rk_combine<T>(nc, c, x0, x1, x2);
}

// halo ring cell h (0 .. nhalo-1): a periodic halo cell takes the
// tendency of its wrapped interior source, a closed-direction one ds = 0
template <class T, int M>
__device__ __forceinline__ void transport_halo(const TransportArgs<T>& a, int64_t h) {
  if (h >= a.nhalo) return;
  const int P = a.P;
  const int NXP = a.nx + 6;
  int j, i;
  if (h < 6 * (int64_t)NXP) {
    int row = (int)(h / NXP);
    i = (int)(h - (int64_t)row * NXP);
    j = row < 3 ? row : a.ny + row;  // rows 0..2, ny+3..ny+5
  } else {
    int64_t u = h - 6 * (int64_t)NXP;
    j = 3 + (int)(u / 6);
    int c = (int)(u % 6);
    i = c < 3 ? c : a.nx + c;
  }
  int js = j, is = i;
  if (a.xper) is = (i < 3) ? i + a.nx : (i >= a.nx + 3 ? i - a.nx : i);
  if (a.yper) js = (j < 3) ? j + a.ny : (j >= a.ny + 3 ? j - a.ny : j);
  T ds = T(0);
  if (js >= 3 && js < 3 + a.ny && is >= 3 && is < 3 + a.nx) {
    int64_t ks = (int64_t)js * P + is;
    T dfx = xface_global<T, M>(a, js, is + 1) - xface_global<T, M>(a, js, is);
    T dfy = yface_global<T, M>(a, js + 1, is) - yface_global<T, M>(a, js, is);
    ds = ((-dfx) - dfy) * T(a.msk[ks]);
  }
  int64_t k = (int64_t)j * P + i;
  stage_update<T>(a, k, ds, a.q[k]);
}

// Tile of workgroup b (counted after the halo-ring blocks, whose number is
// a multiple of 8).  Workgroups are dispatched round-robin over the 8 XCDs,
// workgroup b to XCD b % 8; with nt % 8 == 0 workgroup b takes tile
// (b % 8) * (nt / 8) + b / 8, so each XCD runs one contiguous band of
// tile rows and the halo rows and boundary lines that neighbouring tiles
// share are fetched once into that XCD's L2.
__device__ __forceinline__ int tile_of(int b, int nt) {
  return (nt & 7) ? b : (b & 7) * (nt >> 3) + (b >> 3);
}

// the stage's dt into a.dtv at a transport kernel's start (all 256 threads):
// read from a.dt, or formed from CFL partials (then workgroup 0 stores it
// in dt_w); workgroup 0 also copies it to dt_log
template <class T>
__device__ __forceinline__ void stage_dt(TransportArgs<T>& a) {
  __shared__ double shd[4];
  if (a.cpart) {
    a.dtv = cfl_step256(a.cpart, a.cnb, a.ccfl, a.cdtmax, shd);
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.dt_w = a.dtv;
  } else {
    a.dtv = a.dt ? *a.dt : 0.0;
  }
  if (a.dt_log && blockIdx.x == 0 && threadIdx.x == 0) *a.dt_log = a.dtv;
}

// B: the batched instantiation (a = tab[blockIdx.y]: one member of a
// recorded member batch, launch_rec)
template <class T, int M, int TY, int TX, bool B = false>
__global__ __launch_bounds__(256) void k_transport(TransportArgs<T> a,
                                                   const TransportArgs<T>* tab) {
  if (B) a = tab[blockIdx.y];
  constexpr int LW = TX + 6;
  __shared__ T sq[(TY + 6) * LW];
  __shared__ T sfx[TY * (TX + 1)];
  __shared__ T sfy[(TY + 1) * TX];
  const int t = threadIdx.x;
  const int P = a.P;
  stage_dt(a);

  // the halo-ring blocks come first: they start with the tiles instead of
  // trailing them
  if ((int)blockIdx.x >= a.nhb) {
    const int tile = tile_of((int)blockIdx.x - a.nhb, a.ntiles);
    const int by = tile / a.tiles_x, bx = tile - by * a.tiles_x;
    const int j0 = 3 + by * TY, i0 = 3 + bx * TX;
    const int NYP = a.ny + 6, NXP = a.nx + 6;
    // Every global load of the tile is issued here, before any is used, so
    // that their latencies overlap (no load waits on another load).
    constexpr int NQ = ((TY + 6) * LW + 255) / 256;
    constexpr int NFX = (TY * (TX + 1) + 255) / 256;
    constexpr int NFY = ((TY + 1) * TX + 255) / 256;
    constexpr int NCL = (TY * TX + 255) / 256;
    T qv[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) {
      const int k = t + i * 256;
      const int r = k / LW, c = k - r * LW;
      const int gj = j0 - 3 + r, gi = i0 - 3 + c;
      const bool ok = k < (TY + 6) * LW && gj < NYP && gi < NXP;
      qv[i] = ok ? a.q[(int64_t)gj * P + gi] : T(0);
    }
    int8_t ox[NFX], oy[NFY];
    T ux[NFX], uy[NFY];
#pragma unroll
    for (int i = 0; i < NFX; i++) {
      const int f = t + i * 256;
      const int r = f / (TX + 1), c = f - r * (TX + 1);
      const int gj = j0 + r, gi = i0 + c;
      const bool ok = f < TY * (TX + 1) && gj < 3 + a.ny && gi <= 3 + a.nx;
      const int64_t k = ok ? (int64_t)gj * P + gi : 0;
      ox[i] = ok ? a.ocx[k] : int8_t(0);
      ux[i] = ok ? ux_at(a, gj, gi) : T(0);
    }
#pragma unroll
    for (int i = 0; i < NFY; i++) {
      const int f = t + i * 256;
      const int r = f / TX, c = f - r * TX;
      const int gj = j0 + r, gi = i0 + c;
      const bool ok = f < (TY + 1) * TX && gj <= 3 + a.ny && gi < 3 + a.nx;
      const int64_t k = ok ? (int64_t)gj * P + gi : 0;
      oy[i] = ok ? a.ocy[k] : int8_t(0);
      uy[i] = ok ? uy_at(a, gj, gi) : T(0);
    }
    // the stage's earlier tendencies, as many as its coefficients need
    double cf[3] = {0.0, 0.0, 0.0};
    const int nc = a.integ < 0 ? 0 : rk_coefs(a.integ, a.stage, a.dtv, cf);
    int8_t mv[NCL];
    T d0[NCL], d1[NCL];
#pragma unroll
    for (int i = 0; i < NCL; i++) {
      const int e = t + i * 256;
      const int r = e / TX, c = e - r * TX;
      const int gj = j0 + r, gi = i0 + c;
      const bool ok = e < TY * TX && gj < 3 + a.ny && gi < 3 + a.nx;
      const int64_t k = ok ? (int64_t)gj * P + gi : 0;
      mv[i] = ok ? a.msk[k] : int8_t(0);
      d0[i] = (ok && nc > 1) ? a.ds0[k] : T(0);
      d1[i] = (ok && nc > 2) ? a.ds1[k] : T(0);
    }
#pragma unroll
    for (int i = 0; i < NQ; i++) {
      const int k = t + i * 256;
      if (k < (TY + 6) * LW) sq[k] = qv[i];
    }
    __syncthreads();
    // x faces: rows j0..j0+TY-1, faces i0..i0+TX
#pragma unroll
    for (int i = 0; i < NFX; i++) {
      const int f = t + i * 256;
      if (f >= TY * (TX + 1)) continue;
      const int r = f / (TX + 1), c = f - r * (TX + 1);
      T v = T(0);
      const int o = ox[i];
      if (o > 0) {
        const T u = ux[i];
        const T* sp = sq + (r + 3) * LW + c;  // sp[0] = q[gi-3]
        v = flux_by_order<T, M>(o, u, sp[0], sp[1], sp[2], sp[3], sp[4], sp[5]) * u;
      }
      sfx[f] = v;
    }
    // y faces: faces j0..j0+TY, columns i0..i0+TX-1
#pragma unroll
    for (int i = 0; i < NFY; i++) {
      const int f = t + i * 256;
      if (f >= (TY + 1) * TX) continue;
      const int r = f / TX, c = f - r * TX;
      T v = T(0);
      const int o = oy[i];
      if (o > 0) {
        const T u = uy[i];
        const T* sp = sq + r * LW + (c + 3);  // sp[0] = q[gj-3]
        v = flux_by_order<T, M>(o, u, sp[0], sp[LW], sp[2 * LW], sp[3 * LW],
                                sp[4 * LW], sp[5 * LW]) *
            u;
      }
      sfy[f] = v;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NCL; i++) {
      const int e = t + i * 256;
      const int r = e / TX, c = e - r * TX;
      const int gj = j0 + r, gi = i0 + c;
      if (e >= TY * TX || gj >= 3 + a.ny || gi >= 3 + a.nx) continue;
      const int64_t k = (int64_t)gj * P + gi;
      const T dfx = sfx[r * (TX + 1) + c + 1] - sfx[r * (TX + 1) + c];
      const T dfy = sfy[(r + 1) * TX + c] - sfy[r * TX + c];
      const T ds = ((-dfx) - dfy) * T(mv[i]);
      if (a.dso) a.dso[k] = ds;
      if (nc > 0) {
        const T x0 = (nc == 1) ? ds : d0[i];
        const T x1 = (nc == 2) ? ds : (nc > 2 ? d1[i] : T(0));
        const T x2 = (nc == 3) ? ds : T(0);
        a.qo[k] = sq[(r + 3) * LW + c + 3] + rk_combine<T>(nc, cf, x0, x1, x2);
      }
    }
    return;
  }

  transport_halo<T, M>(a, (int64_t)blockIdx.x * 256 + t);
}

// The same stage for tiles that divide the interior (ny % TY == 0,
// nx % TX == 0, TX a power of two dividing 256): every tile is whole, so
// no load or store needs a bounds test, and thread t owns column t % TX of
// rows t / TX + (256 / TX) i, whose offsets step by a constant.  Only the
// tile's last x-face column and y-face row may lie on the wrapped periodic
// edge (ux_at / uy_at there).  Per face and per cell the arithmetic is
// that of k_transport.
template <class T, int M, int TY, int TX, bool B = false>
__global__ __launch_bounds__(256) void k_transport_full(TransportArgs<T> a,
                                                        const TransportArgs<T>* tab) {
  if (B) a = tab[blockIdx.y];
  constexpr int LW = TX + 6;
  constexpr int RP = 256 / TX;             // rows per pass of the workgroup
  constexpr int NQR = TY + 6;              // q rows of the tile
  constexpr int NQ = (NQR + RP - 1) / RP;  // q rows per thread
  constexpr int NR = TY / RP;              // cell (and face) rows per thread
  static_assert(TX <= 256 && (TX & (TX - 1)) == 0 && TY % RP == 0, "tile shape");
  __shared__ T sq[NQR * LW];
  __shared__ T sfx[TY * (TX + 1)];
  __shared__ T sfy[(TY + 1) * TX];
  __shared__ double shd[8];
  const int t = threadIdx.x;
  if ((int)blockIdx.x < a.nhb) {  // halo-ring blocks first, as in k_transport
    stage_dt(a);
    transport_halo<T, M>(a, (int64_t)blockIdx.x * 256 + t);
    return;
  }
  const uint32_t P = (uint32_t)a.P;
  const int tile = tile_of((int)blockIdx.x - a.nhb, a.ntiles);
  const int by = tile / a.tiles_x, bx = tile - by * a.tiles_x;
  const int j0 = 3 + by * TY, i0 = 3 + bx * TX;
  const int c = t & (TX - 1), r0 = t / TX;
  const uint32_t oc = (uint32_t)(j0 + r0) * P + (uint32_t)(i0 + c);  // cell (r0, c)
  constexpr uint32_t one = 1;
  // every global load first (see k_transport)
  T qv[NQ];
#pragma unroll
  for (int i = 0; i < NQ; i++) {
    const int rr = r0 + RP * i;
    qv[i] = (NQR % RP == 0 || rr < NQR) ? a.q[oc - 3 * P + (uint32_t)(RP * i) * P] : T(0);
  }
  const bool hq = t < 6 * NQR;  // q halo columns: 3 left, 3 right of the tile
  const int hr = t / 6, hcc = t - 6 * (t / 6);
  const int hc = hcc < 3 ? hcc : TX + hcc;  // LDS column
  T qh = T(0);
  if (hq) qh = a.q[(uint32_t)(j0 - 3 + hr) * P + (uint32_t)(i0 - 3 + hc)];
  // the stage's dt: read, or formed from the CFL partials, whose loads go
  // out with the tile's; the 4 wave maxima meet at the tile's first barrier
  double pax = 0.0, pay = 0.0;
  if (a.cpart)
    cfl_partial_max(a.cpart, a.cnb, pax, pay);
  else
    a.dtv = a.dt ? *a.dt : 0.0;
  int8_t ox[NR], oy[NR], mv[NR];
  T ux[NR], uy[NR], d0[NR], d1[NR];
  double cf[3] = {0.0, 0.0, 0.0};
  // (the number of terms does not depend on dt; the coefficients follow)
  const int nc = a.integ < 0 ? 0 : rk_coefs(a.integ, a.stage, 0.0, cf);
#pragma unroll
  for (int i = 0; i < NR; i++) {
    const uint32_t k = oc + (uint32_t)(RP * i) * P;
    ox[i] = a.ocx[k];
    oy[i] = a.ocy[k];
    mv[i] = a.msk[k];
    if (a.psi) {
      ux[i] = ((-(a.psi[k + P] - a.psi[k])) * T(a.pmx[k])) * T(a.psx);
      uy[i] = ((a.psi[k + one] - a.psi[k]) * T(a.pmy[k])) * T(a.psy);
    } else if (a.uscale) {
      ux[i] = a.Ux[k] * T(a.usx);
      uy[i] = a.Uy[k] * T(a.usy);
    } else {
      ux[i] = a.Ux[k];
      uy[i] = a.Uy[k];
    }
    d0[i] = nc > 1 ? a.ds0[k] : T(0);
    d1[i] = nc > 2 ? a.ds1[k] : T(0);
  }
  // the tile's last x-face column (rows t < TY) and y-face row (columns t < TX)
  int8_t oxe = 0, oye = 0;
  T uxe = T(0), uye = T(0);
  if (t < TY) {
    oxe = a.ocx[(uint32_t)(j0 + t) * P + (uint32_t)(i0 + TX)];
    uxe = ux_at(a, j0 + t, i0 + TX);
  }
  if (t < TX) {
    oye = a.ocy[(uint32_t)(j0 + TY) * P + (uint32_t)(i0 + t)];
    uye = uy_at(a, j0 + TY, i0 + t);
  }
#pragma unroll
  for (int i = 0; i < NQ; i++) {
    const int rr = r0 + RP * i;
    if (NQR % RP == 0 || rr < NQR) sq[rr * LW + c + 3] = qv[i];
  }
  if (hq) sq[hr * LW + hc] = qh;
  if (a.cpart) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      pax = nanmax(pax, __shfl_xor(pax, o, 64));
      pay = nanmax(pay, __shfl_xor(pay, o, 64));
    }
    if ((t & 63) == 0) {
      shd[t >> 6] = pax;
      shd[4 + (t >> 6)] = pay;
    }
  }
  __syncthreads();
  if (a.cpart) {
    a.dtv = cfl_of(nanmax(nanmax(shd[0], shd[1]), nanmax(shd[2], shd[3])),
                   nanmax(nanmax(shd[4], shd[5]), nanmax(shd[6], shd[7])), a.ccfl, a.cdtmax);
    if (blockIdx.x == 0 && t == 0) *a.dt_w = a.dtv;
  }
  if (a.dt_log && blockIdx.x == 0 && t == 0) *a.dt_log = a.dtv;
  if (nc > 0) rk_coefs(a.integ, a.stage, a.dtv, cf);
#pragma unroll
  for (int i = 0; i < NR; i++) {
    const int r = r0 + RP * i;
    T v = T(0);
    const int o = ox[i];
    if (o > 0) {
      const T u = ux[i];
      const T* sp = sq + (r + 3) * LW + c;  // sp[0] = q[gi-3]
      v = flux_by_order<T, M>(o, u, sp[0], sp[1], sp[2], sp[3], sp[4], sp[5]) * u;
    }
    sfx[r * (TX + 1) + c] = v;
    T w = T(0);
    const int oo = oy[i];
    if (oo > 0) {
      const T u = uy[i];
      const T* sp = sq + r * LW + (c + 3);  // sp[0] = q[gj-3]
      w = flux_by_order<T, M>(oo, u, sp[0], sp[LW], sp[2 * LW], sp[3 * LW], sp[4 * LW],
                              sp[5 * LW]) *
          u;
    }
    sfy[r * TX + c] = w;
  }
  if (t < TY) {
    T v = T(0);
    if (oxe > 0) {
      const T* sp = sq + (t + 3) * LW + TX;
      v = flux_by_order<T, M>(oxe, uxe, sp[0], sp[1], sp[2], sp[3], sp[4], sp[5]) * uxe;
    }
    sfx[t * (TX + 1) + TX] = v;
  }
  if (t < TX) {
    T w = T(0);
    if (oye > 0) {
      const T* sp = sq + TY * LW + (t + 3);
      w = flux_by_order<T, M>(oye, uye, sp[0], sp[LW], sp[2 * LW], sp[3 * LW], sp[4 * LW],
                              sp[5 * LW]) *
          uye;
    }
    sfy[TY * TX + t] = w;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NR; i++) {
    const int r = r0 + RP * i;
    const uint32_t k = oc + (uint32_t)(RP * i) * P;
    const T dfx = sfx[r * (TX + 1) + c + 1] - sfx[r * (TX + 1) + c];
    const T dfy = sfy[(r + 1) * TX + c] - sfy[r * TX + c];
    const T ds = ((-dfx) - dfy) * T(mv[i]);
    if (a.dso) a.dso[k] = ds;
    if (nc > 0) {
      const T x0 = (nc == 1) ? ds : d0[i];
      const T x1 = (nc == 2) ? ds : (nc > 2 ? d1[i] : T(0));
      const T x2 = (nc == 3) ? ds : T(0);
      a.qo[k] = sq[(r + 3) * LW + c + 3] + rk_combine<T>(nc, cf, x0, x1, x2);
    }
  }
}

template <class T, int M, int TY, int TX>
void launch_transport_tile(const TransportArgs<T>& a0, hipStream_t s) {
  TransportArgs<T> a = a0;
  a.tiles_x = (a.nx + TX - 1) / TX;
  int tiles_y = (a.ny + TY - 1) / TY;
  a.ntiles = a.tiles_x * tiles_y;
  a.nhalo = 6 * (int64_t)(a.nx + 6) + 6 * (int64_t)a.ny;
  // a multiple of 8 halo-ring blocks (the surplus returns at once): tile
  // workgroup b then runs on XCD b % 8 (tile_of)
  a.nhb = (int)(((a.nhalo + 255) / 256 + 7) & ~(int64_t)7);
  int nb = a.ntiles + a.nhb;
  // whole tiles, and every padded offset within 32 bits
  const bool whole = a.ny % TY == 0 && a.nx % TX == 0 &&
                     (int64_t)(a.ny + 6) * (int64_t)a.P < ((int64_t)1 << 31);
  if (whole)
    launch_rec(k_transport_full<T, M, TY, TX, false>, k_transport_full<T, M, TY, TX, true>,
               dim3(nb), dim3(256), 0, s, a);
  else
    launch_rec(k_transport<T, M, TY, TX, false>, k_transport<T, M, TY, TX, true>, dim3(nb),
               dim3(256), 0, s, a);
}

// tile shape by code (f2d_grid.tile): the arithmetic per cell is the same
// for every shape, only the work split and the halo re-reads change
template <class T, int M>
void launch_transport(const TransportArgs<T>& a, int tile, hipStream_t s) {
  switch (tile) {
    case 1: launch_transport_tile<T, M, 8, 64>(a, s); break;
    case 2: launch_transport_tile<T, M, 16, 32>(a, s); break;
    case 3: launch_transport_tile<T, M, 32, 32>(a, s); break;
    case 4: launch_transport_tile<T, M, 8, 128>(a, s); break;
    default: launch_transport_tile<T, M, 16, 64>(a, s); break;
  }
}

template <class T>
__global__ void k_rk_addto(T* __restrict__ y, const T* __restrict__ x0,
                           const T* __restrict__ x1, const T* __restrict__ x2,
                           int64_t n, int integ, int stage,
                           const double* __restrict__ dt) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double c[3];
  int nc = rk_coefs(integ, stage, *dt, c);
  T a0 = x0[k];
  T a1 = nc > 1 ? x1[k] : T(0);
  T a2 = nc > 2 ? x2[k] : T(0);
  y[k] = y[k] + rk_combine<T>(nc, c, a0, a1, a2);
}

// y += ((c0 x0 + c1 x1) + c2 x2) with given coefficients, c_i times *dt when
// bit i of dtmask is set (the nc first terms)
// (y may be one of the x: no restrict; each point is read before written)
template <class T>
__global__ void k_addto(T* y, const T* x0, const T* x1, const T* x2, int64_t n, int nc,
                        double c0, double c1, double c2, const double* __restrict__ dt,
                        int dtmask) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double c[3] = {c0, c1, c2};
  if (dt) {
    const double d = *dt;
    for (int i = 0; i < 3; i++)
      if ((dtmask >> i) & 1) c[i] = c[i] * d;
  }
  T a0 = x0[k];
  T a1 = nc > 1 ? x1[k] : T(0);
  T a2 = nc > 2 ? x2[k] : T(0);
  y[k] = y[k] + rk_combine<T>(nc, c, a0, a1, a2);
}

// ---------------------------------------------------------------------------
// centres -> vertices, operators.py:126-133
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_c2v(const T* __restrict__ c, T* __restrict__ v,
                      const int8_t* __restrict__ mv, int NY, int NX,
                      int addto) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)NY * NX) return;
  int j = (int)(k / NX), i = (int)(k - (int64_t)j * NX);
  T val;
  if (j > 0 && i > 0) {
    T s = T(0.25) * (((c[k - NX - 1] + c[k - 1]) + c[k - NX]) + c[k]);
    val = addto ? v[k] + s : s;
  } else {
    val = v[k];
  }
  v[k] = mv[k] ? val : T(0);
}

// ---------------------------------------------------------------------------
// perpgrad + fill (+ CFL max), operators.py:144-149, equations.py:79-84.
// Every entry of Ux, Uy is written: the rows/columns numpy's slices do not
// reach keep their old value (times the contravariant factor); with fill,
// periodic halo entries take the value of their wrapped source.
// ---------------------------------------------------------------------------
template <class T>
struct PerpArgs {
  const T* psi;
  T* Ux;
  T* Uy;
  const int8_t* mx;
  const int8_t* my;
  int ny, nx;  // interior
  int xper, yper, fill;
  double sx, sy;
  double* partials;  // nullable: 2 * nblocks block maxima of |Ux|, |Uy|
};

constexpr int PG_PTS = 4;  // points per thread of k_perpgrad

template <class T>
__global__ __launch_bounds__(256) void k_perpgrad(PerpArgs<T> a) {
  __shared__ double sh[4];
  const int NY = a.ny + 6, NX = a.nx + 6;
  const int64_t n = (int64_t)NY * NX;
  // chunk of this workgroup: XCD b % 8 takes a contiguous run of chunks
  // (the grid is a multiple of 8), so the row below a chunk, which its
  // x-face differences read, is mostly in the same L2
  const int G = gridDim.x, b = blockIdx.x;
  const int64_t chunk = (b & 7) * (G >> 3) + (b >> 3);
  double mxv = 0.0, myv = 0.0;
  // (j, i) of the first point by one division; each later point is 256
  // further on, and NX > 256 (nx >= 251) keeps it within the next row
  int j = 0, i = 0;
  {
    const int64_t k0 = chunk * PG_PTS * 256 + threadIdx.x;
    if (n < ((int64_t)1 << 31)) {
      j = (int)((uint32_t)k0 / (uint32_t)NX);
      i = (int)((uint32_t)k0 - (uint32_t)j * (uint32_t)NX);
    } else {
      j = (int)(k0 / NX);
      i = (int)(k0 - (int64_t)j * NX);
    }
  }
  for (int r = 0; r < PG_PTS; r++) {
    int64_t k = (chunk * PG_PTS + r) * 256 + threadIdx.x;
    if (r > 0) {
      if (NX > 256) {
        i += 256;
        if (i >= NX) {
          i -= NX;
          j++;
        }
      } else if (n < ((int64_t)1 << 31)) {
        j = (int)((uint32_t)k / (uint32_t)NX);
        i = (int)((uint32_t)k - (uint32_t)j * (uint32_t)NX);
      } else {
        j = (int)(k / NX);
        i = (int)(k - (int64_t)j * NX);
      }
    }
    if (k >= n) break;
    int js = j, is = i;
    if (a.fill) {
      if (a.xper) is = (i < 3) ? i + a.nx : (i >= a.nx + 3 ? i - a.nx : i);
      if (a.yper) js = (j < 3) ? j + a.ny : (j >= a.ny + 3 ? j - a.ny : j);
    }
    int64_t ks = (int64_t)js * NX + is;
    // The last row of Ux / column of Uy is outside numpy's slices: it keeps
    // its value (0 unless written by the caller), scaled like the rest, and
    // its periodic images take the scaled value.  Such an entry is scaled in
    // place by its own thread, which also writes its images; an image's
    // thread leaves it alone (reading the source there would race with the
    // source's own update).  Images equal their sources, so the maxima of
    // the sources are those of every entry.
    if (js < NY - 1 || ks == k) {
      T ux = (js < NY - 1) ? (-(a.psi[ks + NX] - a.psi[ks])) * T(a.mx[ks]) : a.Ux[k];
      ux = ux * T(a.sx);
      a.Ux[k] = ux;
      if (js == NY - 1 && a.fill && a.xper) {
        if (i >= 3 && i < 6) a.Ux[k + a.nx] = ux;
        if (i >= a.nx && i < a.nx + 3) a.Ux[k - a.nx] = ux;
      }
      mxv = nanmax(mxv, fabs((double)ux));
    }
    if (is < NX - 1 || ks == k) {
      T uy = (is < NX - 1) ? (a.psi[ks + 1] - a.psi[ks]) * T(a.my[ks]) : a.Uy[k];
      uy = uy * T(a.sy);
      a.Uy[k] = uy;
      if (is == NX - 1 && a.fill && a.yper) {
        if (j >= 3 && j < 6) a.Uy[k + (int64_t)a.ny * NX] = uy;
        if (j >= a.ny && j < a.ny + 3) a.Uy[k - (int64_t)a.ny * NX] = uy;
      }
      myv = nanmax(myv, fabs((double)uy));
    }
  }
  if (!a.partials) return;
  const double bx = block_nanmax256(mxv, sh);
  const double by = block_nanmax256(myv, sh);
  if (threadIdx.x == 0) {
    a.partials[blockIdx.x] = bx;
    a.partials[gridDim.x + blockIdx.x] = by;
  }
}

// one block: maxima of the partials -> maxabs[2] and the CFL time step
// dt = min(cfl / (max|Ux| + max|Uy| + 1e-99), dtmax)   (model.py:84-87)
__global__ __launch_bounds__(256) void k_dt_fin(const double* part, int nb,
                                                double* maxabs, double* dt,
                                                double cfl, double dtmax) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  double ax = 0.0, ay = 0.0;
  for (int b = t; b < nb; b += 256) {
    ax = nanmax(ax, part[b]);
    ay = nanmax(ay, part[nb + b]);
  }
  sh[t] = ax;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = nanmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  ax = sh[0];
  __syncthreads();
  sh[t] = ay;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = nanmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  ay = sh[0];
  if (t == 0) {
    if (maxabs) {
      maxabs[0] = ax;
      maxabs[1] = ay;
    }
    if (dt) {
      double v = cfl / ((ax + ay) + 1e-99);
      dt[0] = (dtmax < v) ? dtmax : v;  // Python min(v, dtmax)
    }
  }
}

// fill_halo_array (meshes.py:135-143) + y extension: halo entries take
// their wrapped source; entries of closed directions are left alone.
template <class T>
__global__ void k_fill(T* __restrict__ a, int ny, int nx, int xper, int yper) {
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

// f2d_cfl_dt with its block maxima in the caller's array (one per member of
// a recorded member batch): the same maxima, the same dt
struct CflArgs {
  const void* Ux;
  const void* Uy;
  int64_t n;
  double* part;  // 2 nb entries
  int nb;
  double cfl, dtmax;
  double* dt;
};

template <class T, bool B = false>
__global__ __launch_bounds__(256) void k_absmax_r(CflArgs a, const CflArgs* tab) {
  if (B) a = tab[blockIdx.y];
  __shared__ double sh[4];
  const T* __restrict__ Ux = (const T*)a.Ux;
  const T* __restrict__ Uy = (const T*)a.Uy;
  double mx = 0.0, my = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < a.n;
       k += (int64_t)a.nb * 256) {
    mx = nanmax(mx, fabs((double)Ux[k]));
    my = nanmax(my, fabs((double)Uy[k]));
  }
  mx = block_nanmax256(mx, sh);
  my = block_nanmax256(my, sh);
  if (threadIdx.x == 0) {
    a.part[blockIdx.x] = mx;
    a.part[a.nb + blockIdx.x] = my;
  }
}

template <bool B = false>
__global__ __launch_bounds__(256) void k_dt_fin_r(CflArgs a, const CflArgs* tab) {
  if (B) a = tab[blockIdx.y];
  __shared__ double sh[4];
  const double dt = cfl_step256(a.part, a.nb, a.cfl, a.dtmax, sh);
  if (threadIdx.x == 0) *a.dt = dt;
}

template <class T>
__global__ __launch_bounds__(256) void k_absmax(const T* __restrict__ Ux,
                                                const T* __restrict__ Uy,
                                                int64_t n, double* part) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  double mx = 0.0, my = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * 256 + t; k < n;
       k += (int64_t)gridDim.x * 256) {
    mx = nanmax(mx, fabs((double)Ux[k]));
    my = nanmax(my, fabs((double)Uy[k]));
  }
  sh[t] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = nanmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  mx = sh[0];
  __syncthreads();
  sh[t] = my;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = nanmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  if (t == 0) {
    part[blockIdx.x] = mx;
    part[gridDim.x + blockIdx.x] = sh[0];
  }
}

__global__ void k_set_scalar(double* d, double v) { *d = v; }
__global__ void k_copy_f64(const double* __restrict__ src, double* __restrict__ dst,
                           int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Scratch of the grid reductions of this file: partial maxima written by
// one kernel and read by the one-block finaliser that follows it on the same
// stream.  One buffer per device by default (calls on different streams of
// one device must then not overlap); a stream registered with
// f2d_reserve_stream_scratch gets its own, so that work on concurrent
// streams (graph branches) does not share it.
struct ReduceScratch {
  double* partials = nullptr;
  int cap = 0;
};
static ReduceScratch g_red[64];
static std::map<std::pair<int, void*>, ReduceScratch> g_red_stream;
static std::mutex g_red_mu;

static void red_grow(ReduceScratch& r, int nblocks) {
  if (r.cap >= nblocks) return;
  if (r.partials) {
    F2D_HIP(hipDeviceSynchronize());
    F2D_HIP(hipFree(r.partials));
  }
  int cap = nblocks < 8192 ? 8192 : nblocks;
  F2D_HIP(hipMalloc((void**)&r.partials, sizeof(double) * 2 * cap));
  r.cap = cap;
}

static double* red_scratch(int nblocks, void* stream) {
  int dev = 0;
  F2D_HIP(hipGetDevice(&dev));
  F2D_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
  std::lock_guard<std::mutex> lock(g_red_mu);
  auto it = g_red_stream.find({dev, stream});
  ReduceScratch& r = it != g_red_stream.end() ? it->second : g_red[dev];
  red_grow(r, nblocks);
  return r.partials;
}

// Allocation-free variant of the perpgrad + CFL step: the block maxima go
// to `scratch`, a fixed array in each GPU's memory, instead of the
// reduction scratch above (which the library allocates on first use).  In
// host code `scratch` names a host-side stand-in: its device address on the
// current GPU comes from hipGetSymbolAddress, looked up once per device on
// the first call there and kept in g_scratch_dev.  Every call on one GPU
// shares the array: calls that overlap on one GPU (concurrent streams)
// race.  k_perpgrad writes every partial that k_dt_fin reads, so nothing
// needs clearing between calls.
constexpr int NOALLOC_MAX_BLOCKS = 8192;  // 2 x 8192 doubles = 128 KiB
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_BLOCKS];
static double* g_scratch_dev[64];
static std::mutex g_scratch_mu;

static double* noalloc_scratch() {
  int dev = 0;
  F2D_HIP(hipGetDevice(&dev));
  F2D_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    F2D_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    F2D_REQUIRE(p != nullptr, "no device address for the reduction scratch");
    g_scratch_dev[dev] = (double*)p;
  }
  return g_scratch_dev[dev];
}

// workgroups of k_perpgrad over n padded points: a multiple of 8 (its chunk
// order); the surplus chunks are empty and report maxima of 0
static int perpgrad_blocks(int64_t n) { return (grid1d(n, 256 * PG_PTS) + 7) & ~7; }

}  // namespace f2d

using namespace f2d;

#define F2D_DTYPE_DISPATCH(dtype, ...)              \
  do {                                              \
    if ((dtype) == F2D_F64) {                       \
      using T = double;                             \
      __VA_ARGS__;                                  \
    } else if ((dtype) == F2D_F32) {                \
      using T = float;                              \
      __VA_ARGS__;                                  \
    } else {                                        \
      throw ::f2d::Error{"unknown dtype"};          \
    }                                               \
  } while (0)

static void check_grid(const f2d_grid* g) {
  F2D_REQUIRE(g != nullptr, "grid descriptor is NULL");
  F2D_REQUIRE(g->halowidth == 3, "halowidth must be 3");
  F2D_REQUIRE(g->ny > 0 && g->nx > 0, "grid is empty");
  F2D_REQUIRE(g->ny < (1 << 24) && g->nx < (1 << 24), "grid too large");
}

static void check_method(int m) {
  F2D_REQUIRE(m >= F2D_WENO && m <= F2D_CWENO, "unknown flux method");
}

static void launch_perpgrad(int dtype, const f2d_grid* g, const void* psi, void* Ux,
                            void* Uy, const int8_t* mskx, const int8_t* msky,
                            double sx, double sy, int fill, double* part, int nb,
                            double* maxabs, double* dt, double cfl, double dtmax,
                            void* stream) {
  F2D_DTYPE_DISPATCH(dtype, {
    PerpArgs<T> a{};
    a.psi = (const T*)psi;
    a.Ux = (T*)Ux;
    a.Uy = (T*)Uy;
    a.mx = mskx;
    a.my = msky;
    a.ny = (int)g->ny;
    a.nx = (int)g->nx;
    a.xper = g->xperiodic;
    a.yper = g->yperiodic;
    a.fill = fill;
    a.sx = sx;
    a.sy = sy;
    a.partials = part;
    k_perpgrad<T><<<nb, 256, 0, S(stream)>>>(a);
  });
  if (part && (maxabs || dt))
    k_dt_fin<<<1, 256, 0, S(stream)>>>(part, nb, maxabs, dt, cfl, dtmax);
}

extern "C" {

int f2d_compflux(int dtype, void* flx, const void* U, const void* q,
                 const int8_t* o, int64_t n, int64_t s, int method,
                 void* stream) {
  return guarded([&] {
    check_method(method);
    F2D_REQUIRE(n > 0 && s != 0, "f2d_compflux: bad size/shift");
    F2D_DTYPE_DISPATCH(dtype, k_compflux<T><<<grid1d(n, 256), 256, 0,
                                              S(stream)>>>(
                                  (T*)flx, (const T*)U, (const T*)q, o, n, s,
                                  method));
  });
}

int f2d_vortexforce(int dtype, void* du, const void* V, const void* q,
                    const int8_t* o, int64_t n, int64_t s, int64_t s2,
                    int sign, int method, void* stream) {
  return guarded([&] {
    check_method(method);
    F2D_REQUIRE(n > 0 && s != 0, "f2d_vortexforce: bad size/shift");
    F2D_DTYPE_DISPATCH(dtype, k_vortexforce<T><<<grid1d(n, 256), 256, 0,
                                                 S(stream)>>>(
                                  (T*)du, (const T*)V, (const T*)q, o, n, s,
                                  s2, sign, method));
  });
}

int f2d_innerproduct(int dtype, void* ke, const void* U, const void* q,
                     const int8_t* o, int64_t n, int64_t s, int method,
                     void* stream) {
  return guarded([&] {
    check_method(method);
    F2D_REQUIRE(n > 0 && s != 0, "f2d_innerproduct: bad size/shift");
    F2D_DTYPE_DISPATCH(dtype, k_innerproduct<T><<<grid1d(n, 256), 256, 0,
                                                  S(stream)>>>(
                                  (T*)ke, (const T*)U, (const T*)q, o, n, s,
                                  method));
  });
}

static int transport_stage_impl(int dtype, const f2d_grid* g, int method,
                                const void* q_in, void* q_out, const void* Ux,
                                const void* Uy, const void* psi, const int8_t* mskx,
                                const int8_t* msky, double sx, double sy, int fill,
                                const int8_t* ocx, const int8_t* ocy, const int8_t* msk,
                                const void* ds_prev0, const void* ds_prev1, void* ds_out,
                                int integrator, int stage, const double* dt, void* stream,
                                double* dt_log = nullptr, bool from_u = false,
                                const double* cpart = nullptr, int64_t npart = 0,
                                double ccfl = 0.0, double cdtmax = 0.0) {
  return guarded_rec([&] {
    check_grid(g);
    check_method(method);
    F2D_REQUIRE(!dt_log || dt, "dt_log needs dt");
    if (cpart) {
      F2D_REQUIRE(integrator >= 0, "a stage dt from CFL partials needs an integrator");
      F2D_REQUIRE(npart == 2 * (int64_t)perpgrad_blocks((g->ny + 6) * (g->nx + 6)),
                  "npart must be f2d_cfl_partials() of this grid");
    }
    if (psi) {
      F2D_REQUIRE(mskx && msky, "mskx and msky are required with psi");
    } else {
      F2D_REQUIRE(Ux && Uy, "Ux and Uy are required");
    }
    if (integrator >= 0) {
      F2D_REQUIRE(integrator <= F2D_ENRK3, "unknown integrator");
      int nst = integrator == F2D_EF ? 1 : 3;
      F2D_REQUIRE(stage >= 0 && stage < nst, "bad stage");
      F2D_REQUIRE(q_out && q_out != q_in, "q_out must be a distinct array");
      F2D_REQUIRE(dt, "dt is NULL");
      double c[3];
      int nc = rk_coefs(integrator, stage, 0.0, c);
      F2D_REQUIRE(nc < 2 || ds_prev0, "ds_prev0 is required at this stage");
      F2D_REQUIRE(nc < 3 || ds_prev1, "ds_prev1 is required at this stage");
    } else {
      F2D_REQUIRE(ds_out, "rhs-only call needs ds_out");
    }
    F2D_DTYPE_DISPATCH(dtype, {
      TransportArgs<T> a{};
      a.q = (const T*)q_in;
      a.qo = (T*)q_out;
      a.Ux = (const T*)Ux;
      a.Uy = (const T*)Uy;
      a.psi = (const T*)psi;
      a.pmx = mskx;
      a.pmy = msky;
      a.psx = sx;
      a.psy = sy;
      a.pfill = fill;
      a.uscale = from_u;
      a.usx = 1.0 / (g->dx * g->dx);  // sharp's 1/dx**2, 1/dy**2
      a.usy = 1.0 / (g->dy * g->dy);
      a.ocx = ocx;
      a.ocy = ocy;
      a.msk = msk;
      a.ds0 = (const T*)ds_prev0;
      a.ds1 = (const T*)ds_prev1;
      a.dso = (T*)ds_out;
      a.dt_log = dt_log;
      a.dt = dt;
      a.cpart = cpart;
      a.cnb = (int)(npart / 2);
      a.ccfl = ccfl;
      a.cdtmax = cdtmax;
      a.dt_w = const_cast<double*>(dt);
      a.integ = integrator;
      a.stage = stage;
      a.ny = (int)g->ny;
      a.nx = (int)g->nx;
      a.P = (int)g->nx + 6;
      a.xper = g->xperiodic;
      a.yper = g->yperiodic;
      switch (method) {
        case F2D_WENO: launch_transport<T, F2D_WENO>(a, g->tile, S(stream)); break;
        case F2D_UPWIND: launch_transport<T, F2D_UPWIND>(a, g->tile, S(stream)); break;
        case F2D_CENTERED: launch_transport<T, F2D_CENTERED>(a, g->tile, S(stream)); break;
        default: launch_transport<T, F2D_CWENO>(a, g->tile, S(stream)); break;
      }
    });
  });
}

int f2d_transport_stage(int dtype, const f2d_grid* g, int method,
                        const void* q_in, void* q_out, const void* Ux,
                        const void* Uy, const int8_t* ocx, const int8_t* ocy,
                        const int8_t* msk, const void* ds_prev0,
                        const void* ds_prev1, void* ds_out, int integrator,
                        int stage, const double* dt, void* stream) {
  return transport_stage_impl(dtype, g, method, q_in, q_out, Ux, Uy, nullptr, nullptr,
                              nullptr, 0.0, 0.0, 0, ocx, ocy, msk, ds_prev0, ds_prev1,
                              ds_out, integrator, stage, dt, stream);
}

int f2d_transport_stage_psi(int dtype, const f2d_grid* g, int method, const void* q_in,
                            void* q_out, const void* psi, const int8_t* mskx,
                            const int8_t* msky, double sx, double sy, int fill,
                            const int8_t* ocx, const int8_t* ocy, const int8_t* msk,
                            const void* ds_prev0, const void* ds_prev1, void* ds_out,
                            int integrator, int stage, const double* dt, void* stream) {
  // a NULL psi reaches the Ux/Uy check of the common path and is refused
  return transport_stage_impl(dtype, g, method, q_in, q_out, nullptr, nullptr, psi, mskx,
                              msky, sx, sy, fill, ocx, ocy, msk, ds_prev0, ds_prev1, ds_out,
                              integrator, stage, dt, stream);
}

int f2d_transport_stage_u(int dtype, const f2d_grid* g, int method, const void* q_in,
                          void* q_out, const void* ux, const void* uy, const int8_t* ocx,
                          const int8_t* ocy, const int8_t* msk, const void* ds_prev0,
                          const void* ds_prev1, void* ds_out, int integrator, int stage,
                          const double* dt, void* stream) {
  return transport_stage_impl(dtype, g, method, q_in, q_out, ux, uy, nullptr, nullptr,
                              nullptr, 0.0, 0.0, 0, ocx, ocy, msk, ds_prev0, ds_prev1,
                              ds_out, integrator, stage, dt, stream, nullptr, true);
}

int f2d_transport_stage_log(int dtype, const f2d_grid* g, int method, const void* q_in,
                            void* q_out, const void* Ux, const void* Uy, const int8_t* ocx,
                            const int8_t* ocy, const int8_t* msk, const void* ds_prev0,
                            const void* ds_prev1, void* ds_out, int integrator, int stage,
                            const double* dt, double* dt_log, void* stream) {
  return transport_stage_impl(dtype, g, method, q_in, q_out, Ux, Uy, nullptr, nullptr,
                              nullptr, 0.0, 0.0, 0, ocx, ocy, msk, ds_prev0, ds_prev1,
                              ds_out, integrator, stage, dt, stream, dt_log);
}

int f2d_transport_stage_cfl(int dtype, const f2d_grid* g, int method, const void* q_in,
                            void* q_out, const void* Ux, const void* Uy, const int8_t* ocx,
                            const int8_t* ocy, const int8_t* msk, const void* ds_prev0,
                            const void* ds_prev1, void* ds_out, int integrator, int stage,
                            double* dt, const double* part, int64_t npart, double cfl,
                            double dtmax, double* dt_log, void* stream) {
  if (!part) {
    set_error("f2d_transport_stage_cfl: part is NULL");
    return -1;
  }
  return transport_stage_impl(dtype, g, method, q_in, q_out, Ux, Uy, nullptr, nullptr,
                              nullptr, 0.0, 0.0, 0, ocx, ocy, msk, ds_prev0, ds_prev1,
                              ds_out, integrator, stage, dt, stream, dt_log, false, part,
                              npart, cfl, dtmax);
}

int f2d_rk_addto(int dtype, void* y, const void* x0, const void* x1,
                 const void* x2, int64_t n, int integrator, int stage,
                 const double* dt, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(integrator >= 0 && integrator <= F2D_ENRK3, "bad integrator");
    F2D_DTYPE_DISPATCH(dtype, k_rk_addto<T><<<grid1d(n, 256), 256, 0,
                                              S(stream)>>>(
                                  (T*)y, (const T*)x0, (const T*)x1,
                                  (const T*)x2, n, integrator, stage, dt));
  });
}

int f2d_addto(int dtype, void* y, const void* x0, const void* x1, const void* x2,
              int64_t n, int nterms, double c0, double c1, double c2, const double* dt,
              int dtmask, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(nterms >= 1 && nterms <= 3, "nterms must be 1, 2 or 3");
    F2D_REQUIRE(y && x0 && (nterms < 2 || x1) && (nterms < 3 || x2), "NULL argument");
    F2D_REQUIRE(!dtmask || dt, "dtmask needs dt");
    F2D_DTYPE_DISPATCH(dtype, k_addto<T><<<grid1d(n, 256), 256, 0, S(stream)>>>(
                                  (T*)y, (const T*)x0, (const T*)x1, (const T*)x2, n, nterms,
                                  c0, c1, c2, dtmask ? dt : nullptr, dtmask));
  });
}

int f2d_c2v(int dtype, const f2d_grid* g, const void* c, void* v,
            const int8_t* mskv, int addto, void* stream) {
  return guarded([&] {
    check_grid(g);
    int NY = (int)g->ny + 6, NX = (int)g->nx + 6;
    F2D_DTYPE_DISPATCH(dtype, k_c2v<T><<<grid1d((int64_t)NY * NX, 256), 256,
                                         0, S(stream)>>>(
                                  (const T*)c, (T*)v, mskv, NY, NX, addto));
  });
}

int f2d_perpgrad_ex(int dtype, const f2d_grid* g, const void* psi, void* Ux,
                    void* Uy, const int8_t* mskx, const int8_t* msky,
                    double sx, double sy, int fill, double* maxabs,
                    double* dt, double cfl, double dtmax, void* stream) {
  return guarded([&] {
    check_grid(g);
    int64_t n = (g->ny + 6) * (g->nx + 6);
    int nb = perpgrad_blocks(n);
    double* part = (maxabs || dt) ? red_scratch(nb, stream) : nullptr;
    launch_perpgrad(dtype, g, psi, Ux, Uy, mskx, msky, sx, sy, fill, part, nb, maxabs,
                    dt, cfl, dtmax, stream);
  });
}

int f2d_perpgrad_cfl_noalloc(int dtype, const f2d_grid* g, const void* psi, void* Ux,
                             void* Uy, const int8_t* mskx, const int8_t* msky,
                             double sx, double sy, int fill, double* maxabs,
                             double* dt, double cfl, double dtmax, void* stream) {
  return guarded([&] {
    check_grid(g);
    F2D_REQUIRE(maxabs || dt, "f2d_perpgrad_cfl_noalloc: maxabs and dt are both NULL "
                              "(use f2d_perpgrad_ex for the gradient alone)");
    int64_t n = (g->ny + 6) * (g->nx + 6);
    int nb = perpgrad_blocks(n);
    if (nb > NOALLOC_MAX_BLOCKS)
      throw Error{"f2d_perpgrad_cfl_noalloc: " + std::to_string(n) +
                     " padded points need " + std::to_string(nb) +
                     " reduction blocks; the fixed scratch holds " +
                     std::to_string(NOALLOC_MAX_BLOCKS) + " (at most " +
                     std::to_string((int64_t)NOALLOC_MAX_BLOCKS * 256 * PG_PTS) +
                     " padded points, e.g. 2890 x 2890 cells)"};
    double* part = noalloc_scratch();
    launch_perpgrad(dtype, g, psi, Ux, Uy, mskx, msky, sx, sy, fill, part, nb, maxabs,
                    dt, cfl, dtmax, stream);
  });
}

int f2d_cfl_partials(const f2d_grid* g, int64_t* npart) {
  return guarded([&] {
    check_grid(g);
    F2D_REQUIRE(npart != nullptr, "npart is NULL");
    *npart = 2 * (int64_t)perpgrad_blocks((g->ny + 6) * (g->nx + 6));
  });
}

int f2d_perpgrad_part(int dtype, const f2d_grid* g, const void* psi, void* Ux, void* Uy,
                      const int8_t* mskx, const int8_t* msky, double sx, double sy, int fill,
                      double* part, int64_t npart, void* stream) {
  return guarded([&] {
    check_grid(g);
    const int nb = perpgrad_blocks((g->ny + 6) * (g->nx + 6));
    F2D_REQUIRE(part != nullptr, "part is NULL");
    F2D_REQUIRE(npart == 2 * (int64_t)nb, "npart must be f2d_cfl_partials() of this grid");
    launch_perpgrad(dtype, g, psi, Ux, Uy, mskx, msky, sx, sy, fill, part, nb, nullptr,
                    nullptr, 0.0, 0.0, stream);
  });
}

int f2d_cfl_finish(const double* part, int64_t npart, double* maxabs, double* dt, double cfl,
                   double dtmax, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(part != nullptr, "part is NULL");
    F2D_REQUIRE(npart > 0 && npart % 2 == 0 && npart / 2 < (1 << 30),
                "npart must be f2d_cfl_partials() of the grid");
    F2D_REQUIRE(maxabs || dt, "maxabs and dt are both NULL");
    k_dt_fin<<<1, 256, 0, S(stream)>>>(part, (int)(npart / 2), maxabs, dt, cfl, dtmax);
  });
}

int f2d_perpgrad(int dtype, const f2d_grid* g, const void* psi, void* Ux,
                 void* Uy, const int8_t* mskx, const int8_t* msky,
                 int contravariant, double* maxabs, void* stream) {
  double sx = 1.0, sy = 1.0;
  if (contravariant) {
    sx = 1.0 / (g->dy * g->dy);
    sy = 1.0 / (g->dx * g->dx);
  }
  return f2d_perpgrad_ex(dtype, g, psi, Ux, Uy, mskx, msky, sx, sy, 1, maxabs,
                         nullptr, 0.0, 0.0, stream);
}

int f2d_fill_halo(int dtype, const f2d_grid* g, void* a, void* stream) {
  return guarded([&] {
    check_grid(g);
    if (!g->xperiodic && !g->yperiodic) return;
    int64_t n = (g->ny + 6) * (g->nx + 6);
    F2D_DTYPE_DISPATCH(dtype, k_fill<T><<<grid1d(n, 256), 256, 0,
                                          S(stream)>>>(
                                  (T*)a, (int)g->ny, (int)g->nx,
                                  g->xperiodic, g->yperiodic));
  });
}

int f2d_cfl_dt(int dtype, const void* Ux, const void* Uy, int64_t n,
               double cfl, double dtmax, double* dt, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(n > 0, "f2d_cfl_dt: empty");
    F2D_REQUIRE(dt, "dt is NULL");
    int nb = grid1d(n, 256);
    if (nb > 1024) nb = 1024;
    double* part = red_scratch(nb, stream);
    F2D_DTYPE_DISPATCH(dtype, k_absmax<T><<<nb, 256, 0, S(stream)>>>(
                                  (const T*)Ux, (const T*)Uy, n, part));
    k_dt_fin<<<1, 256, 0, S(stream)>>>(part, nb, nullptr, dt, cfl, dtmax);
  });
}

int f2d_cfl_dt_part(int dtype, const void* Ux, const void* Uy, int64_t n, double cfl,
                    double dtmax, double* part, int64_t npart, double* dt, void* stream) {
  return guarded_rec([&] {
    F2D_REQUIRE(n > 0, "f2d_cfl_dt_part: empty");
    F2D_REQUIRE(Ux && Uy && dt && part, "f2d_cfl_dt_part: NULL argument");
    int nb = grid1d(n, 256);
    if (nb > 1024) nb = 1024;
    F2D_REQUIRE(npart >= 2 * (int64_t)nb, "f2d_cfl_dt_part: npart must be at least 2048 "
                                          "(or 2 ceil(n / 256))");
    CflArgs a{Ux, Uy, n, part, nb, cfl, dtmax, dt};
    F2D_DTYPE_DISPATCH(dtype, launch_rec(k_absmax_r<T, false>, k_absmax_r<T, true>, dim3(nb),
                                         dim3(256), 0, S(stream), a));
    launch_rec(k_dt_fin_r<false>, k_dt_fin_r<true>, dim3(1), dim3(256), 0, S(stream), a);
  });
}

int f2d_set_scalar(double* dst, double value, void* stream) {
  return guarded([&] { k_set_scalar<<<1, 1, 0, S(stream)>>>(dst, value); });
}

int f2d_reserve_stream_scratch(void* stream) {
  return guarded([&] {
    int dev = 0;
    F2D_HIP(hipGetDevice(&dev));
    F2D_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
    std::lock_guard<std::mutex> lock(g_red_mu);
    red_grow(g_red_stream[{dev, stream}], 8192);
  });
}

int f2d_copy_f64(double* dst, const double* src, int n, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(n >= 0 && (n == 0 || (dst && src)), "bad copy arguments");
    if (n > 0) k_copy_f64<<<(n + 255) / 256, 256, 0, S(stream)>>>(src, dst, n);
  });
}

}  // extern "C"
