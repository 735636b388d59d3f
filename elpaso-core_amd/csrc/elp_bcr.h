// Direct solve of a model's structure block by block cyclic reduction.
//
// The free structure rows (field 0) are ordered by a node coordinate (the
// axis giving the smallest half-bandwidth b), which makes the structure
// block banded; cut into blocks of m = b rows it is block tridiagonal,
//   A = tridiag(C_i, D_i, C_{i+1}^T),  C_i = A[block i, block i-1],
// complex symmetric (A = ks K + ms M with complex ks, ms).  Cyclic reduction
// eliminates every other block per level:
//   W_k = C_k D_{k-1}^-1,  V_k = C_{k+1}^T D_{k+1}^-1,
//   D'_k = D_k - W_k C_k^T - V_k C_{k+1},  C'_k = -W_k C_{k-1},
// so a solve is ~2 log2(N) levels of batched dense products over all the
// blocks of a level (no long sequential chain), on the matrix cores through
// rocBLAS; the block inverses of the eliminated blocks are LU with partial
// pivoting (rocSOLVER).  Used as the structure part of the coupled
// preconditioner (a thin clamped plate is too ill-conditioned for smoothers
// or multigrid); the factorization is refreshed per window of frequencies.
#pragma once

#include <complex>
#include <memory>

#include "elp_internal.h"

namespace elp {

struct BcrPlan;
template <typename T> struct BcrFactor;

// nullptr when the model has no free structure rows
std::unique_ptr<BcrPlan, void (*)(BcrPlan *)> bcr_plan(const Model &m, const std::vector<int32_t> *iperm);
int bcr_block_size(const BcrPlan &p);
int bcr_blocks(const BcrPlan &p);
int64_t bcr_rows(const BcrPlan &p);
int64_t bcr_factor_bytes(const BcrPlan &p, int elem_bytes);

template <typename T>
std::unique_ptr<BcrFactor<T>, void (*)(BcrFactor<T> *)> bcr_new_factor();

// factor A_ss = ks K + ms M (the model's FP64 K, M) over the plan's rows
template <typename T>
void bcr_factor(const Model &m, BcrPlan &p, std::complex<double> ks, std::complex<double> ms, BcrFactor<T> &f,
                hipStream_t s);

// z[row][c] = (A_ss^-1 r)[row][c] on the plan's rows for the B columns of
// the [row][B] vectors r, z (solver row order); other rows of z untouched
// (model_order: r, z in the model's row order instead of the FP32
// operator's locality order)
template <typename T, typename VR>
void bcr_apply(BcrPlan &p, const BcrFactor<T> &f, int B, const VR *r, VR *z, hipStream_t s, bool model_order);

}  // namespace elp
