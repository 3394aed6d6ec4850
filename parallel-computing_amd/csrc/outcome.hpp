// Host-side decision after each FP64 residual check of a solve: the
// convergence test of PETSc's KSPConvergedDefault as the reference's KSP
// uses it (matrix_free/iteration_number_fempc_mf.cpp:219-224: rtol, abstol,
// divtol 1e4, maxit), with the residual norms squared.  Plain C++, no HIP:
// compiled into the library and into the host-only sanitizer build.
#pragma once

#include <cmath>

#include "iga_mf.h"

namespace igai {

// Returned by outer_outcome while the solve should go on.
constexpr int OUTCOME_CONTINUE = -1;

// rr = ||f - A u||^2, ff = ||f||^2, tol2 = max(rtol ||f||, abstol)^2.
// at_maxit: the iteration count reached maxit.  In order: a non-finite f is
// IGA_ERR_NOT_FINITE (an input error, found before any iteration); a NaN
// residual or one above divtol ||f|| is IGA_ERR_DIVERGED; a residual at or
// below the tolerance is IGA_OK; at maxit IGA_ERR_NOT_CONVERGED; otherwise
// OUTCOME_CONTINUE.
inline int outer_outcome(double rr, double ff, double tol2, double divtol, bool at_maxit) {
  if (!std::isfinite(ff)) return IGA_ERR_NOT_FINITE;
  if (!(rr == rr) || rr > divtol * divtol * ff) return IGA_ERR_DIVERGED;
  if (rr <= tol2) return IGA_OK;
  if (at_maxit) return IGA_ERR_NOT_CONVERGED;
  return OUTCOME_CONTINUE;
}

}  // namespace igai
