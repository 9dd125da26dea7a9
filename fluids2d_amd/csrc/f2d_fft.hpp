// f2d_fft.hpp -- direct spectral solve of the 5-point Poisson/Helmholtz
// operator of fully wet domains (internal interface of the library): doubly
// periodic as below, closed directions through odd (vertex) or even (cell)
// extensions on the doubled periodic grid.
//
// When every interior vertex is an unknown and both directions are periodic
// (elliptic.py:114-195 with the y-periodic extension), the operator
//   (A x)_P = cx (x_E + x_W) + cy (x_N + x_S) - (2 (cx + cy) + maindiag) x_P
// is diagonalised by the 2-D DFT, eigenvalues
//   lambda(kx, ky) = 2 cx (cos(2 pi kx / nx) - 1) + 2 cy (cos(2 pi ky / ny) - 1)
//                    - maindiag,
// so x = IDFT(DFT(b) / lambda), the k = 0 mode of the singular operator set
// to zero (the zero-mean solution the iterative solver also returns).
#pragma once
#include "f2d_common.hpp"

namespace f2d {

struct FftSolver;

// nx, ny powers of two >= 8, every vertex strictly inside the domain
// unknown: all nx x ny of a doubly periodic one; in a closed direction the
// boundary lines are Dirichlet (nx - 1 unknowns) and the problem is solved
// as its odd extension on the periodic doubled grid (2 nx <= 4096 then)
bool fft_applicable(int64_t ny, int64_t nx, int xper, int yper, int64_t nwet);

// the Neumann cell operator (elliptic.py:186-190) of an all-wet domain:
// every cell unknown, nx, ny powers of two >= 8 (<= 2048 when closed); a
// closed direction is solved as the even extension about its walls
bool fft_applicable_cells(int64_t ny, int64_t nx, int xper, int yper, int64_t nwet);

// the transforms of a row and of a column fit in a workgroup's LDS
// (FP64: transformed sizes up to 2048; FP32: up to 4096)
bool fft_fits(int64_t ny, int64_t nx, int xper, int yper, bool f64);

// cells: the Neumann cell operator (even extensions), else the vertex
// operator (odd extensions of closed directions)
FftSolver* fft_create(int64_t ny, int64_t nx, double dx, double dy,
                      double maindiag, bool f64, bool f32, int xper, int yper,
                      bool cells = false);
void fft_destroy(FftSolver* f);

// x = A^-1 b on the padded (ny+6) x (nx+6) layout, halo filled.  gmode 0:
// b given at the vertices; 1: b = centerstovertices(omega) of the cell
// field passed as `b`.  FP64 throughout.  With `stats` (8 doubles, device)
// the relative residual of the result is measured and recorded.
void fft_solve_f64(FftSolver* f, const double* b, int gmode, double* x,
                   double rtol, bool measure, double* stats, hipStream_t s);

// FP32 transforms with FP64 iterative refinement: x64 = S32(b), then
// `nref` times r = b - A x64 (FP64), x64 += S32(r); x = (float) x64.
// warm (from the handle's second solve on): no cold solve, the refinement
// starts from the caller's x (its interior; the halo is not read):
// x64 = x + S32(b - A x), then nref - 1 further refinements.
void fft_solve_f32_refine(FftSolver* f, const float* b, int gmode, float* x,
                          int nref, bool warm, double rtol, bool measure,
                          double* stats, hipStream_t s);

// Batched FP32 spectral solves of same-size members (blockIdx.y = member):
// each member's right-hand side b (FP32, padded, gmode 0), spectrum S and
// solution x; `f` is one member's plan (fft_same_plan holds for all).
// Masks of a spectral solve inside the capacitance solve (all optional):
// the forward row pass reads b only where in[k] (0 elsewhere) -- with xr,
// the FP64 residual b - A x of the operator masked to in (x float, read at
// in only) -- and writes that right-hand side to out (b's type); the
// inverse row pass writes (plain) or adds (accumulating) its result to x
// only where res[k], without periodic halo images.
struct FftMask {
  const int8_t* in = nullptr;
  void* out = nullptr;
  const float* xr = nullptr;
  const int8_t* res = nullptr;
};
struct FftBatchPtrs {
  const void* b;
  void* S;
  void* x;
  FftMask m;
};
// acc: the inverse row pass adds its result to x in FP64 (x = (float)(x + e))
void fft_solve_f32_batch(FftSolver* f, int nm, const FftBatchPtrs* table, hipStream_t s,
                         bool acc = false);
// one plain FP32 spectral solve with masks (x = the result, gmode as b's)
void fft_solve_f32_masked(FftSolver* f, const float* b, int gmode, float* x, const FftMask& m,
                          hipStream_t s, bool acc = false);
bool fft_same_plan(const FftSolver* f, const FftSolver* g);
float2* fft_spectrum32(FftSolver* f);

// timing hook (see f2d_poisson_bench_kernel)
double fft_bench(FftSolver* f, bool f64, int kind, int reps, hipStream_t s,
                 double* bytes);

}  // namespace f2d
