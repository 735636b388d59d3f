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

// Grouped serving: requests of one precision and slot count advance
// together, one COCG iteration of each per step, with their operator, dot
// products and vector updates fused into single launches (bit-identical to
// running each request alone).
struct GroupItem {
  int64_t tag = 0;  // the caller's request id
  virtual ~GroupItem() = default;
};
struct GroupDriver {
  virtual ~GroupDriver() = default;
  // start a prepared request (on the driver's stream)
  virtual void add(std::unique_ptr<GroupItem> it) = 0;
  virtual bool empty() const = 0;
  // one iteration of every request; requests whose sweep ended are moved to
  // `finished` (their results are in the arrays given to group_prepare)
  virtual void step(std::vector<std::unique_ptr<GroupItem>> &finished) = 0;
};
// Set-up of a request on `setup` (solver state, work space).  *key: the
// group it runs in, by precision, slot count and size class (-1: it runs
// alone through sweep(), nullptr returned).
std::unique_ptr<GroupItem> group_prepare(Model &m, int64_t nf, const double *freqs, std::complex<double> Z,
                                         const elp_solve_opts &o, double *X, double *relres, int32_t *iters,
                                         hipStream_t setup, int *key);
std::unique_ptr<GroupDriver> group_driver(int key, Ctx *ctx, hipStream_t s);

}  // namespace elp
