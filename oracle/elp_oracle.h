/*
 * elp_oracle.h -- CPU restatement of elPaSo-Core's FE assembly and harmonic
 * system set-up, in plain C (FP64, complex where elPaSo is complex).
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle the GPU results are compared
 * with: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links, loads or calls it.
 *
 * Every function cites the elPaSo source it restates (paths relative to
 * elpasoCore/source).  Loop orders and the order of floating-point operations
 * follow the reference so that results agree with elPaSo to rounding.
 *
 * Pinning (what checks each function against something other than itself):
 *   gauss tables      -> literal table + sum of weights + polynomial exactness
 *   shape functions   -> reference golden vectors (test/shape/ headers) + partition
 *                        of unity, Kronecker property, derivative sums
 *   jacobian          -> affine closed form (det J = V/8)
 *   fluid K/M         -> rectangular closed form (tensor products of 1-D
 *                        matrices), K*1 = 0, sum(M) = V/c^2
 *   brick K/M         -> 6 rigid-body modes in the null space, rank 18,
 *                        cube closed form for M, sum(M block) = rho*V
 *   Hex20 fluid       -> no elPaSo element exists; pinned by the same
 *                        invariants and quadratic-field patch test
 *   surface integral  -> planar rectangle closed form (M_x (x) M_y)
 *   FSI coupling      -> planar face closed form, row sums = area * normal
 *   global assembly   -> dense brute force on tiny meshes
 */
#ifndef ELP_ORACLE_H
#define ELP_ORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* derivative selector, elPaSo eDerivative order: N_fun, N_xi, N_eta, N_zeta */
enum { ORC_N = 0, ORC_NXI = 1, ORC_NETA = 2, ORC_NZETA = 3 };

/* Gauss-Legendre tables (misc/gausspoints.h:28-45, gausspoints.cpp:34-48) */
double orc_gauss_point_1d(int level, int nr);
double orc_gauss_weight_1d(int level, int nr);
/* 3D: x fastest (gausspoints.cpp:76-87, 184-195) */
void orc_gauss_point_3d(int level, int nr, double gp[3]);
double orc_gauss_weight_3d(int level, int nr);
/* 2D: y = nr % level, x = nr / level (gausspoints.cpp:165-174, 266-274) */
void orc_gauss_point_2d(int level, int nr, double gp[3]);
double orc_gauss_weight_2d(int level, int nr);

/* shape functions: hex nnod in {8,20}, quad nnod in {4,8}
 * (shape/shapefunctions.cpp:110-195 Quad4, 640-955 Hex20, 957-1096 Hex8) */
double orc_shape_hex(int nnod, int k, int deriv, const double gp[3]);
double orc_shape_quad(int nnod, int k, int deriv, const double gp[3]);

/* cShapeHex8 tables, cArray3d layout data[gp*nf*nn + deriv*nn + k]
 * (shape/shapehex8.cpp:49-90, misc/array3d.h:55-62) */
void orc_hex8_table(double N[256]);
void orc_hex8_face_table(double Nface[48]);
/* face node indices, 0-based face id (shape/shapehex8.cpp:119-146) */
void orc_hex8_face_nodes(int face, int idx[4]);
/* Hex20 faces (8 nodes: 4 corners counter-clockwise seen from outside, then
 * the 4 midsides in Quad8 order).  No elPaSo counterpart (Brick20/Fluid20 are
 * not in elPaSo-Core); same face numbering/normals as the Hex8 faces. */
void orc_hex20_face_nodes(int face, int idx[8]);

/* Jacobian at one Gauss point of a hex element (element/elementfem.cpp:158-210).
 * xyz: nnod x 3.  Returns detJ; J and invJ row-major 3x3. */
double orc_jacobian3d(int nnod, int ngp_level, int gpnr, const double *xyz,
                      double J[9], double invJ[9]);

/* element matrices, nnod in {8,20}; 2x2x2 Gauss for 8 nodes, 3x3x3 for 20 */
/* cElementFluid3d::assembleMassMatrix (element/fluid/elementfluid3d.cpp:54-98) */
void orc_fluid_mass(int nnod, const double *xyz, double c, double *MM);
/* cElementFluid3d::assembleStiffnessMatrix (elementfluid3d.cpp:100-167) */
void orc_fluid_stiffness(int nnod, const double *xyz, double *KM);
/* cElementFluid3d::evaluateSurfaceIntegral (elementfluid3d.cpp:169-234);
 * nnod_face 4 (level 2) or 8 (Quad8, level 3). Returns 0 or -1 if detJ <= 0. */
int orc_surface_integral(int nnod_face, const double *xyz_face, double *C);

/* cMaterialStructureIsotrop::setupC / cMaterialStructureVisco::setupC
 * (materialstructureisotrop.cpp:130-147, materialstructurevisco.cpp:349-372) */
void orc_setup_c(double complex E, double nu, double complex C[36]);
/* E_vis (materialstructurevisco.cpp:283-285) */
double complex orc_e_visco(double E, double eta, int type, double omega);
/* cElementStructureBrick::assembleStiffnessMatrix linear path
 * (elementstructurebrick.cpp:405, 470-508) with infam::BT_C_B_wdJ
 * (math/matrix.h:757-790).  KM complex nnod*3 square, ADDED to. */
void orc_brick_stiffness(int nnod, const double *xyz, double complex E,
                         double nu, double complex *KM);
/* cElementStructureBrick::assembleMassMatrix (elementstructurebrick.cpp:511-554) */
void orc_brick_mass(int nnod, const double *xyz, double rho, double *MM);

/* cElementFluid::assembleDynamicStiffnessMatrix (element/fluid/elementfluid.cpp:58-77)
 * EM = (-w^2 M + K + Z_imp)/(rho w^2).  imp_faces: list of 0-based faces with
 * Robin impedance Z (complex), lumped as in the commented formula
 * elementfluid3d.cpp:324-369 when lumped != 0, consistent otherwise. */
void orc_fluid_dynamic(int nnod, const double *xyz, double rho, double c,
                       double omega, int n_imp, const int *imp_faces,
                       double complex Z, int lumped, double complex *EM);
/* cElementStructure::assembleDynamicStiffnessMatrix (elementstructure.cpp:43-61) */
void orc_brick_dynamic(int nnod, const double *xyz, double complex Evis,
                       double nu, double rho, double omega,
                       double complex *EM);

/* cElementInterfaceMindlin::computeCouplingMatrix + transformC
 * (element/interface/elementinterfacemindlin.cpp:80-181): 4-node face, fluid
 * node coordinates xyz (4x3); C is 12 x 4 row-major, ADDED to. */
void orc_coupling_matrix(const double *xyz, double *C);

/* ------------------------------------------------------------------------
 * Global assembly (analysis/analysis.cpp:806-867 pattern, 1238-1246 scatter;
 * analysis/frequency/analysisfrequencybasic.cpp:281-304 element loop)
 * ---------------------------------------------------------------------- */

/* CSR pattern from element DOF cliques (+ diagonal), sorted, 0-based.
 * edofs: n_elem x ndpe global rows.  rowptr must hold n+1 ints.  Call with
 * colind == NULL to get nnz (returned), then again with colind allocated. */
int64_t orc_pattern(int n, int n_elem, int ndpe, const int *edofs,
                    int n_extra, const int *extra_rows, const int *extra_cols,
                    int *rowptr, int *colind);

/* element kinds for the global routines */
enum { ORC_FLUID8 = 1, ORC_BRICK8 = 2, ORC_FLUID20 = 3, ORC_BRICK20 = 4 };

/* Assemble the real K and M of a group of elements into CSR values (ADDED),
 * elements in the given order.  For fluid kinds the element matrices are
 * divided by rho (the frequency-independent part of elPaSo's 1/(rho w^2)
 * scaling) when fluid_rho_scale != 0.  Returns 0, or -(1+e) if element e has
 * detJ <= 0. */
int orc_assemble_km(int kind, int n_elem, const int *conn, const double *xyz,
                    const int *edofs, double p0, double p1, double p2,
                    int fluid_rho_scale, const int *rowptr, const int *colind,
                    double *K, double *M);

/* Assemble elPaSo's dynamic matrix A(w) of a group (complex, ADDED), the
 * per-frequency element re-integration of assembleGlobalTensors.
 * p0..p2: fluid (rho, c, -) / brick (E, nu, rho); Evis for bricks.
 * imp: per element 6-bit mask of Robin faces (may be NULL). */
int orc_assemble_dynamic(int kind, int n_elem, const int *conn,
                         const double *xyz, const int *edofs, double p0,
                         double p1, double p2, double complex Evis,
                         double omega, const unsigned char *imp_mask,
                         double complex Z, int lumped, const int *rowptr,
                         const int *colind, double complex *A);

/* Robin face matrix C_G of a group of faces (real, ADDED): lumped -> row sums
 * on the diagonal, consistent -> full face matrix. face_nodes: nf x nfn global
 * node indices in face order; fdofs: nf x nfn global rows. */
int orc_assemble_face(int nfn, int n_faces, const int *face_nodes,
                      const double *xyz, const int *fdofs, int lumped,
                      const int *rowptr, const int *colind, double *C);

/* FSI coupling (elementinterfacestructural.cpp:46-69, 152-179, 209-237):
 * adds -ori*C into (structure rows, fluid cols) and -ori*C^T into (fluid rows,
 * structure cols). fnodes: n x 4 fluid node indices (geometry). */
int orc_assemble_fsi(int n_if, const int *fnodes, const double *xyz,
                     const int *srows, const int *frows, const double *ori,
                     const int *rowptr, const int *colind, double *S);

/* sparse lookups (binary search in a sorted CSR row); -1 if absent */
int64_t orc_find(const int *rowptr, const int *colind, int r, int c);

#ifdef __cplusplus
}
#endif
#endif
