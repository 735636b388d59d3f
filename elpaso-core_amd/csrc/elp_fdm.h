// Exact solver of the trilinear (Hex8) acoustic operator on a uniform box by
// fast diagonalization: the Hex8 stiffness, mass and z-side Robin matrices of
// a uniform box are Kronecker sums/products of 1-D matrices, so
//   A = (1/rho) [ (Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz)/w^2
//                 - (1 - i beta)/c^2 Mx(x)My(x)Mz + i rho/(w Z) Mx(x)My(x)Ez ]
// is diagonalized in x and y by the M-orthonormal eigenvectors of (Kx, Mx)
// and (Ky, My), leaving one complex tridiagonal system along z per (x-mode,
// y-mode, frequency).  One solve = two dense axis transforms, a batch of
// Thomas solves, and the two inverse transforms.
#pragma once

#include <complex>
#include <memory>

#include "elp_internal.h"

namespace elp {

struct FdmCoef {
  double w2inv;  // 1/w^2
  double2 sM;    // (1 - i beta)/c^2
  double2 tG;    // i rho / (w Z)
  double rho;
};

struct FdmBox;

constexpr int kFdmMaxAxis = 256;  // nodes along x and y of the box lattice

// nullptr when the model is not a uniform fluid box with Robin faces on z
// sides only; with fluid_part, the box describes the fluid rows of a model
// that has other rows too (a Hex8 cavity coupled to a structure)
std::unique_ptr<FdmBox, void (*)(FdmBox *)> fdm_build(const Model &m, bool fluid_part = false);
// the same for a box description bx instead of the model's own; mode 0:
// the model is the box's fluid, 1: its fluid rows are exactly the box's
// lattice (a Hex8 cavity beside a structure), 2: a stand-in box of the
// model's one fluid (the solver maps the fluid rows onto its lattice)
std::unique_ptr<FdmBox, void (*)(FdmBox *)> fdm_build(const Model &m, const BoxInfo &bx, int mode);
int64_t fdm_nodes(const FdmBox &f);
const int *fdm_dims(const FdmBox &f);
double fdm_rho(const FdmBox &f);
double fdm_c(const FdmBox &f);

// x = A^{-1} b for B frequencies ([node][B] complex layout, node = lattice
// index i + px (j + py k)); scratch: 2 vectors of the same size.
void fdm_solve(const FdmBox &f, int B, const FdmCoef *coef, const float2 *b, float2 *x, float2 *s1, float2 *s2,
               hipStream_t s);
void fdm_solve(const FdmBox &f, int B, const FdmCoef *coef, const double2 *b, double2 *x, double2 *s1,
               double2 *s2, hipStream_t s);

}  // namespace elp
