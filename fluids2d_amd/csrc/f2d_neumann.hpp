// f2d_neumann.hpp -- the Neumann ("c", cell-centred) Poisson operator of
// get_Laplacian_sparse (elliptic.py:114-195, neumann branch of :186-190):
// the off-diagonals dy/dx (E/W) and dx/dy (N/S) of the unknown neighbours
// (x-periodic wrap, y-periodic as an extension), the diagonal minus their
// sum minus maindiag.  Solved by Jacobi-preconditioned CG in FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace f2d {

struct NeumannSolver;

// m_host: int8 unknown mask on the padded (ny+6) x (nx+6) grid (interior only)
NeumannSolver* neu_create(const int8_t* m_host, int64_t ny, int64_t nx, int xper,
                          int yper, double dx, double dy, double maindiag);
void neu_destroy(NeumannSolver* n);

// Solve A x = b at the unknowns (b, x of dtype F2D_F64 or F2D_F32, padded;
// x is the initial guess and is written at the unknowns only, then its
// periodic halo images).  The operator is singular without mass term
// (a constant null vector per connected wet region): b is projected onto
// its range (its mean over each region removed, which leaves the
// divergence of a wall-bounded flux unchanged up to rounding) and x keeps
// the constant of its initial guess.  At most 64 disconnected regions.
// sync: check convergence on the host every few iterations and report
// iters / relres; otherwise maxit iterations are enqueued, the ones after
// convergence returning at once.
void neu_solve(NeumannSolver* n, int dtype, const void* b, void* x, double rtol,
               int maxit, bool sync, int* iters, double* relres, hipStream_t s);

double* neu_stats(NeumannSolver* n);  // 8 device doubles, as f2d_poisson_stats
void neu_reset_stats(NeumannSolver* n, hipStream_t s);

}  // namespace f2d
