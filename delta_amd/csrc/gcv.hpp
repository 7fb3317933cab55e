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
};

size_t gcv_scratch_bytes(int p, int k, int nl);
void gcv_solve(const GcvProblem& g, Arena& ar, hipStream_t s);

// Pivoted LDLT fallback for one grid point (Eigen's diagonal-pivoting
// algorithm): theta = A^-1 rhs and edf = tr(A^-1 M).
size_t ldlt_scratch_bytes(int p);
void ldlt_solve_point(const double* M, int ldm, const double* rhs, const double* pen, int k,
                      double lambda, int p, double* theta, double* edf, Arena& ar,
                      hipStream_t s);

}  // namespace dg
