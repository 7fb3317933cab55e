// Batched FP64 GCV solve of the penalised normal equations (gcv.hip).
#pragma once

#include "dg_internal.hpp"

namespace dg {

struct GcvProblem {
  int p, k, rank, nl;       // unknowns, knots per field, penalty rank, grid size
  const double* M;          // device, P x P with leading dimension ldm
  int ldm;
  const double* rhs;        // device, P
  const double* pen;        // device, k x k
  const double* fk;         // device, k x k (first `rank` columns used): Fk Fk^T = pen
  const double* lambdas;    // device, nl
  double* theta;            // device out, nl x P
  double* sq;               // device out, nl: ||L^-1 F||_F^2
  int* fail;                // device out, nl: 1 where the Cholesky failed
  // Optional per-grid-point systems: grid point l uses M + base[l] * m_stride
  // and rhs + base[l] * rhs_stride (device, nl entries; null: one system).
  const int* base = nullptr;
  size_t m_stride = 0, rhs_stride = 0;
  // Replay the launch sequence as a HIP graph, captured on first use for
  // these exact arguments (the IRLS passes and steps of a fit repeat them);
  // DELTA_AMD_GRAPHS=0 launches the kernels one by one instead.
  bool graph = false;
};

size_t gcv_scratch_bytes(int p, int k, int nl);
void gcv_solve(const GcvProblem& g, Arena& ar, hipStream_t s);

// Pivoted LDLT fallback (Eigen's diagonal-pivoting algorithm) for the grid
// points listed in `which` (host indices into lambdas, device), all at once:
// theta[l] = A_l^-1 rhs and edf[l] = tr(A_l^-1 M) (device, indexed by l).
size_t ldlt_scratch_bytes(int p, int nb);
void ldlt_solve_batch(const double* M, int ldm, const double* rhs, const double* pen, int k,
                      const double* lambdas, const std::vector<int>& which, int p, double* theta,
                      double* edf, Arena& ar, hipStream_t s);

}  // namespace dg
