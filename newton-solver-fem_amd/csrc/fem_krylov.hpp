// Krylov methods besides PCG for the step's linear solve (reference
// SolverFEM._make_linear_solver solver.py:1219-1252, which calls Warp's
// warp.optim.linear.bicgstab / cr / gmres with tol = linear_solver_tolerance,
// atol = 0, maxiter, M, warm start dv) and the point-Jacobi preconditioners
// of warp.optim.linear.preconditioner ("diag", "diag_abs", "id";
// solver.py:1155-1156).
//
// The solve works on the step's normalised projected system (the same
// b, warm start and operator PAP + I - P as the persistent PCG) through a
// fixed sequence of kernel launches per iteration: no host synchronisation,
// so a step stays graph-capturable.  Updates are masked by the residual test
// (rr > atol^2) in every step, so running the whole budget gives the result
// of stopping at convergence.  Dot products are summed in fp64 per block and
// then over blocks in a fixed order (deterministic).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fem {

enum KrylovMethod { KR_CG = 0, KR_BICGSTAB = 1, KR_CR = 2, KR_GMRES = 3 };
enum KrylovPrecond { KP_BLOCK = 0, KP_DIAG = 1, KP_DIAG_ABS = 2, KP_ID = 3 };

constexpr int KR_BLOCK = 256;
constexpr int KR_PART = 4;        // partial sums per block and reduction
constexpr int KR_MAX_RESTART = 64;

struct KrylovArgs {
  int P;
  const int* offsets;
  const int* cols;
  const float* vals;               // A (unprojected values)
  const unsigned char* row_fixed;  // operator PAP + I - P over these rows
  const float* M;                  // per-node 3x3 preconditioner (row-major); NULL: identity
  const float* rhs;                // projected right-hand side from the force kernel
  const unsigned* rhs_max;         // max|b| words (float bits) of the single instance
  int rhs_max_slots;
  float* dv;                       // warm start in, increment out
  const float* q_in;
  const float* qd_in;
  float* q_out;                    // NULL: solve only (dv written, no state update)
  float* qd_out;
  const float* anchor;
  float h;
  float tolerance;
  int maxiter;
  int restart;                     // GMRES cycle length (<= KR_MAX_RESTART)
  float* work;                     // krylov_work_floats(P, method, restart)
  double* part;                    // [blocks][KR_PART]
  double* sc;                      // scalars (KR_NSC)
  int* stats;                      // iterations, rr (float bits), atol^2 (float bits)
};

constexpr int KR_NSC = 32 + (KR_MAX_RESTART + 1) * (KR_MAX_RESTART + 1) + 2 * (KR_MAX_RESTART + 1);

// Work vectors (floats), partial-sum stride and partial-sum entries of a solve.
size_t krylov_work_floats(int P, int method, int restart);
inline int krylov_part_stride(int restart) { return restart + 1 > KR_PART ? restart + 1 : KR_PART; }
inline size_t krylov_part_doubles(int P, int restart) {
  return (size_t)((P + KR_BLOCK / 4 - 1) / (KR_BLOCK / 4)) * krylov_part_stride(restart);
}

// The launches of one solve on `st`.
hipError_t krylov_solve(const KrylovArgs& a, int method, hipStream_t st);

// Point-Jacobi preconditioner blocks from A's diagonal blocks: the inverse of
// each diagonal entry (of its magnitude for KP_DIAG_ABS), 1 where it is zero
// or the row stores no diagonal block; identity blocks for KP_ID.
hipError_t krylov_point_jacobi(int P, const int* offsets, const int* diag_rel, const float* vals, int kind, float* out,
                               hipStream_t st);

}  // namespace fem
