// Harmonic sweep solver: fused multi-frequency operator, box / multigrid /
// l1-Jacobi preconditioners, COCG on continuously refilled frequency slots,
// FP64 refinement.
#pragma once

#include <complex>

#include "elp_internal.h"

namespace elp {

void default_opts(elp_solve_opts &o);

// Solve A(w_k) x_k = f for all frequencies (see elp_sweep in elpaso_amd.h);
// returns the number of frequencies whose FP64 relative residual is above
// o.rtol (their best solutions are still returned).
int64_t sweep(Model &m, int64_t nf, const double *freqs, std::complex<double> Z, const elp_solve_opts &o,
           double *X, double *relres, int32_t *iters, elp_sweep_stats *stats, hipStream_t s);

// device bytes the solver needs for `block` frequency slots
int64_t solver_bytes_estimate(const Model &m, int block, int precision);

}  // namespace elp
