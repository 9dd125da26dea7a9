// f2d_capacitance.hpp -- direct solve of the masked vertex operator through
// the rectangle's spectral solve and a capacitance matrix (internal
// interface of the library).
//
// The reference's masked Dirichlet operator (elliptic.py:114-195, "v") on
// the wet unknowns W is the principal submatrix A_R[W, W] of the operator
// A_R of the whole rectangle (every vertex strictly inside it an unknown):
// the same diagonal at every unknown, couplings only between unknown
// neighbours.  With B the rectangle unknowns that are dry but have a wet
// neighbour, x = A_R^-1 (b_W + P_B^T s) solves A_R[W,W] x_W = b_W once
// x_B = 0, i.e. s = M g with g = (A_R^-1 b_W)_B and M = (-C)^-1,
// C = P_B A_R^-1 P_B^T (the capacitance matrix, negative definite).
// A_R^-1 is the spectral solve (f2d_fft); C comes from the Green's function
// of the periodic doubled grid (one spectral solve of a point source, plus
// the odd images of closed directions); M is formed once by a Cholesky
// factorisation and inversion (rocSOLVER).  A solve is then two spectral
// solves and one |B| x |B| matrix-vector product.
#pragma once
#include "f2d_common.hpp"

namespace f2d {

struct CapSolver;

// Number of boundary points B of the unknown mask (padded layout), or -1
// when the method does not apply: the rectangle's spectral solve must apply
// (power-of-two sizes, the FP64 Green's function on the doubled grid within
// the LDS limits), the mask must not cover the whole rectangle, every
// unknown must be a rectangle unknown, and the rectangle operator must be
// non-singular (a closed direction or maindiag > 0).
int64_t cap_boundary(const int8_t* unk, int64_t ny, int64_t nx, int xper, int yper,
                     double maindiag, bool f64_mode);

// f64_mode: FP64 solves (M in FP64); otherwise FP32 spectral solves and an
// FP32 M inside FP64 refinement.  Synchronises the device.
CapSolver* cap_create(const int8_t* unk, int64_t ny, int64_t nx, double dx, double dy,
                      double maindiag, int xper, int yper, bool f64_mode);
void cap_destroy(CapSolver* c);
int64_t cap_size(const CapSolver* c);

// x = A^-1 b at the unknowns (other entries untouched, halo not filled).
// gmode 0: b at the vertices; 1: b = centerstovertices of the cell field.
// measure: the relative residual goes to `stats` (8 doubles, device) as
// the spectral solve records it.
void cap_solve_f64(CapSolver* c, const double* b, int gmode, double* x, bool measure,
                   double* stats, hipStream_t s);
// FP32: passes of  r = b - A x64 (FP64),  x64 += S32(r), S32 the FP32
// capacitance solve; cold: x64 = 0 and nref + 1 passes; warm (from the
// second solve on): x64 = the caller's x and nref passes.  x = (float) x64.
void cap_solve_f32(CapSolver* c, const float* b, int gmode, float* x, int nref, bool warm,
                   bool measure, double* stats, hipStream_t s);

// Batched FP32 solves of same-size members (one launch per kernel of the
// pass for all of them; blockIdx.y = member).  Every member must be an FP32
// capacitance solver of the same rectangle operator (grid, spacings,
// maindiag, periodicity); masks and boundaries may differ.  b[i] and x[i]
// are member i's fixed right-hand side and solution (gmode 0).  A solve
// runs member i's cap_solve_f32 sequence entry for entry, so the results
// are those of the member's own solve, bit for bit.
struct CapBatch;
CapBatch* cap_batch_create(CapSolver* const* c, int nm, const float* const* b,
                           float* const* x);
void cap_batch_destroy(CapBatch* bt);
// whether every member has solved before (a warm batch needs all of them)
bool cap_batch_warm(const CapBatch* bt);
// nref 0: plain FP32 solves (no FP64 refinement)
void cap_batch_solve(CapBatch* bt, int nref, bool warm, hipStream_t s);

}  // namespace f2d
