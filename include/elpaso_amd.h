/*
 * elpaso_amd.h -- C-ABI of the MI355X-native FE assembly and harmonic
 * frequency-sweep library (libelpaso_amd.so).
 *
 * Conventions
 *  - Every function returns 0 on success and a negative ELP_E* code on error;
 *    nothing throws across the ABI.  elp_last_error() gives the message of
 *    the last failing call on a context (thread-local).
 *  - Complex numbers are interleaved (re, im) doubles, the layout of C99
 *    double _Complex, std::complex<double> and PETSc's complex PetscScalar.
 *  - Ids are elPaSo ids.  Rows follow elPaSo's numbering: nodes in ascending
 *    id, DOFs per node by enum value (disp_x1 = 0, disp_x2 = 1, disp_x3 = 2,
 *    fluid = 16), each active DOF takes the next row
 *    (analysis/analysis.cpp:475-520).
 *  - CSR is 0-based, columns sorted within rows, diagonal always present, the
 *    layout PETSc's MatCreateSeqAIJWithArrays takes.
 *  - All device work is on one GPU; matrices stay resident on the device.
 *
 * Reference interfaces each group replaces (paths relative to
 * elpasoCore/source of elPaSo-Core):
 *  - model definition  <- cFemParserInterface::parseInputFile filling cMesh
 *    (misc/parser/femparserinterface.cpp:189-663), cAnalysis::activateDofsAtNodes
 *    (analysis/analysis.cpp:475-520), getLocalSparsityPattern/preallocateMatrix
 *    (analysis.cpp:783-867)
 *  - element operators <- cElementFEM::assembleStiffnessMatrix /
 *    assembleMassMatrix / assembleDynamicStiffnessMatrix
 *    (element/elementfem.h:124-159)
 *  - global assembly   <- cAnalysisFrequencyBasic::assembleGlobalTensors
 *    (analysis/frequency/analysisfrequencybasic.cpp:281-304) with
 *    cAnalysis::insertElementMatrix (analysis.cpp:1238-1246),
 *    cElementInterface::addContributionFrequencyDomain
 *    (element/interface/elementinterface.h:233-234),
 *    cAnalysis::applyNodalLoads / insertBoundaryConditions
 *    (analysis.cpp:165-279, 419-454)
 *  - sweep             <- the frequency loop of cAnalysisFrequencyBasic::FullRun
 *    (analysisfrequencybasic.cpp:74-279) and cAnalysis::solveEquationSystem
 *    (analysis.cpp:1293-1332)
 */
#ifndef ELPASO_AMD_H
#define ELPASO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define ELP_OK 0
#define ELP_EINVAL -1   /* invalid argument / inconsistent model */
#define ELP_EHIP -2     /* HIP runtime error */
#define ELP_EMEM -3     /* device memory budget exceeded */
#define ELP_EDETJ -4    /* element with det J <= 0 */
#define ELP_ESTATE -5   /* call out of order (e.g. assemble before finalize) */
#define ELP_ENOCONV -6  /* solver did not reach the tolerance */

/* element kinds (elPaSo ElementType strings "Fluid8", "Brick8"; the 20-node
 * kinds have no elPaSo-Core element and use the serendipity shape functions of
 * shape/shapefunctions.cpp:640-955) */
#define ELP_FLUID8 1
#define ELP_BRICK8 2
#define ELP_FLUID20 3
#define ELP_BRICK20 4

/* matrices held by an assembled model */
#define ELP_MAT_K 0  /* stiffness; fluid rows divided by rho             */
#define ELP_MAT_M 1  /* mass; fluid rows carry 1/(rho c^2)                */
#define ELP_MAT_C 2  /* Rayleigh damping alpha K + beta M                 */
#define ELP_MAT_G 3  /* Robin face matrix C_Gamma (fluid rows)            */
#define ELP_MAT_S 4  /* FSI coupling with -ori (structure/fluid blocks)   */

/* solver precision / method */
#define ELP_PREC_FP64 0   /* complex FP64 Krylov + FP64 preconditioner, one frequency at a time */
#define ELP_PREC_MIXED 1  /* complex FP32 Krylov/preconditioner, FP64 refinement, frequency blocks */
#define ELP_METHOD_COCG 0 /* the only method: A(w) is complex symmetric */
#define ELP_PC_JACOBI 0   /* l1-Jacobi */
#define ELP_PC_MG 1       /* shifted-Laplacian multigrid (box hierarchy) */

typedef struct elp_ctx elp_ctx;
typedef struct elp_model elp_model;
typedef struct elp_server elp_server;

typedef struct elp_solve_opts {
  int precision;     /* ELP_PREC_*                                        */
  int method;        /* ELP_METHOD_*                                      */
  int precond;       /* ELP_PC_*                                          */
  int block;         /* mixed: frequency slots solved together (1, 4, 8,  */
                     /* 16 or 32; a slot takes the next frequency as soon */
                     /* as its current one is done)                       */
  double rtol;       /* FP64 relative residual target, e.g. 1e-8          */
  int max_iters;     /* Krylov iterations per frequency                   */
  double inner_rtol; /* mixed: FP32 inner reduction per refinement step   */
                     /* (the last step aims just past rtol when it can)   */
  double shift;      /* shifted-Laplacian damping beta (MG); < 0 (the     */
                     /* default) automatic: 0 where the box solve is    */
                     /* exact (Hex8 box, coupled cavity), growing with  */
                     /* k h on Hex20 and stand-in boxes, 0.5 otherwise  */
  int smooth;        /* MG pre/post smoothing sweeps                      */
  int coarse_sweeps; /* MG coarsest-level smoothing sweeps                */
  int mg_levels;     /* MG: maximum number of levels (0 = automatic)      */
  int mg_min_nodes;  /* MG: stop coarsening below this many nodes         */
  int op_tile_rows;  /* 32-slot FP32 operator variant: -1 one row per    */
                     /* wavefront (also for the box cycle's fused post-   */
                     /* smoothing), -2 / -4 rows grouped 2 / 4 to a       */
                     /* wavefront, t (< 1000) LDS tiles of t rows,        */
                     /* 1000 R + t LDS tiles of t rows with rows grouped  */
                     /* R = 2 or 4 to a wavefront (e.g. 4032), 6000 + t   */
                     /* those grouped 4 with each tile's frequencies      */
                     /* split over two workgroups; 0 = chosen at set-up   */
                     /* by timing candidates on the model                 */
  int graph;         /* FP32 sweeps: each Krylov iteration replayed from  */
                     /* a captured HIP graph: 1 on, -1 off, 0 for models  */
                     /* up to 100,000 rows (launch-bound)                 */
} elp_solve_opts;

typedef struct elp_sweep_stats {
  double setup_ms;     /* preconditioner set-up                           */
  double solve_ms;     /* solve wall time                                 */
  int64_t spmv_calls;  /* fine-level fused SpMV launches                  */
  double spmv_ms;      /* time in the fine-level fused SpMV (HIP events)  */
  int64_t spmv_bytes;  /* algorithmic bytes of one fine-level SpMV launch */
  int refinements;     /* FP64 refinement steps (mixed)                   */
  int unconverged;     /* frequencies above rtol (then ELP_ENOCONV)       */
  int refactors;       /* coupled: factorizations of the structure block  */
  double refactor_ms;  /* coupled: time spent in them                     */
  int op_tile_rows;    /* fine-level operator used (the op_tile_rows      */
                       /* encoding), 0 another operator (other slots)     */
  int64_t graph_replays; /* iterations replayed from the captured graph   */
  int box_tile_rows;     /* Hex20 box cycle's post-smoothing: rows per     */
                         /* grouped half-line tile, 0 one row per         */
                         /* wavefront or no fused post-smoothing          */
} elp_sweep_stats;

/* ---- context -------------------------------------------------------------- */
int elp_ctx_create(elp_ctx **out, int device);
void elp_ctx_destroy(elp_ctx *ctx);
const char *elp_last_error(const elp_ctx *ctx);
int elp_default_solve_opts(elp_solve_opts *opts);

/* ---- model definition (mirrors the cMesh the HDF5 parser builds) ---------- */
int elp_model_create(elp_ctx *ctx, elp_model **out);
void elp_model_destroy(elp_model *m);
int elp_model_set_nodes(elp_model *m, int64_t n, const int64_t *ids, const double *xyz);
int elp_model_add_fluid_material(elp_model *m, int64_t id, double rho, double c);
/* eta > 0: viscoelastic, E_vis = E(1 + i eta) (type 1) or E(1 + i w eta) (type 2)
 * (material/structure/isotrop/linear/viscoelastic/materialstructurevisco.cpp:283-285);
 * all viscoelastic materials of one model must share eta and type */
int elp_model_add_structure_material(elp_model *m, int64_t id, double E, double nu,
                                     double rho, double eta, int visco_type);
/* conn: n x npe node ids in elPaSo local node order */
int elp_model_add_elements(elp_model *m, int kind, int64_t material_id, int64_t n,
                           const int64_t *elem_ids, const int64_t *conn);
/* Robin impedance faces (elem id, 1-based face as elPaSo element loads,
 * elementfluid3d.cpp:259).  lumped != 0 puts the face matrix row sums on the
 * diagonal as the commented formula of elementfluid3d.cpp:324-369 does. */
int elp_model_add_impedance_faces(elp_model *m, int64_t n, const int64_t *elem_ids,
                                  const int32_t *faces, double z_re, double z_im,
                                  int lumped);
/* FSI interface Quad4 elements (element/interface/elementinterfacestructural.cpp):
 * fluid_nodes / struct_nodes: n x 4 node ids (geometry from the fluid nodes) */
int elp_model_add_interfaces(elp_model *m, int64_t n, const int64_t *ids, const double *ori,
                             const int64_t *fluid_nodes, const int64_t *struct_nodes);
int elp_model_add_point_source(elp_model *m, int64_t node_id, double Q);
int elp_model_add_point_force(elp_model *m, int64_t node_id, double fx, double fy, double fz);
int elp_model_add_dirichlet(elp_model *m, int64_t node_id, int dof, double re, double im);
int elp_model_set_rayleigh(elp_model *m, double alpha, double beta);
/* Optional: the model is a structured box (origin, edge lengths, elements per
 * direction, order 1 = Hex8 / 2 = Hex20, Robin faces on the box sides in
 * robin_sides: bit k = 0-based Hex8 face numbering -y,+y,-z,+x,+z,-x).
 * Enables the multigrid preconditioner.  For a model with structure rows the
 * box describes its Hex8 fluid cavity (order 1): the sweep then uses the
 * block-diagonal coupled preconditioner (exact box solve of the fluid rows,
 * direct block cyclic reduction of the structure block).  Without it,
 * elp_model_finalize reads the box from the mesh when the fluid elements
 * (one kind) tile a uniform axis-aligned lattice and every Robin face lies on
 * a box side (models read from elPaSo input files included). */
int elp_model_set_box(elp_model *m, const double *origin, const double *length,
                      const int32_t *nel, int order, int robin_sides);
/* The model's box (finalized model): *set = 0 when it has none, *detected =
 * 1 when it was read from the mesh.  Any output pointer may be NULL.  No
 * counterpart in the reference (elPaSo solves every model with its direct
 * solver). */
int elp_model_get_box(const elp_model *m, int *set, int *detected, double *origin, double *length,
                      int32_t *nel, int *order, int *robin_sides);
/* DOF numbering, symbolic phase (sparsity pattern, adjacency), upload.
 * ELP_EINVAL for a node coupled to more than 256 nodes (rows may have any
 * number of columns),
 * ELP_EDETJ (at assembly) for an element with detJ <= 0. */
int elp_model_finalize(elp_model *m);
int elp_model_info(const elp_model *m, int64_t *n_dof, int64_t *nnz, int64_t *device_bytes);
/* /Solution/Maps/mtxDofMap: row -> (node id, dof enum) */
int elp_model_dof_map(const elp_model *m, int64_t *node_ids, int32_t *dofs);

/* ---- assembly -------------------------------------------------------------- */
/* assembles K, M (and C, C_Gamma, S where present) on the device, FP64 */
int elp_assemble(elp_model *m, double *elapsed_ms);
/* run only the volume-element assembly kernels again (timing), in ms of GPU time */
int elp_assemble_timed(elp_model *m, int repeats, double *kernel_ms);
int elp_get_pattern(const elp_model *m, int32_t *rowptr, int32_t *colind);
/* values of one matrix in the full pattern (host copy) */
int elp_get_values(const elp_model *m, int which, double *vals);
/* device pointers of the full-pattern CSR (valid until elp_model_destroy) */
int elp_get_device_csr(const elp_model *m, int which, const int32_t **rowptr,
                       const int32_t **colind, const double **vals);
/* elPaSo's system at one frequency in PETSc AIJ layout (values in the
 * elp_get_pattern order) after nodal loads and Dirichlet rows:
 * A(w) = fluid rows: K/w^2 - M + i/(w Z) C_G ; structure rows:
 * (1 + i eta g) K - w^2 M (+ i w C if Rayleigh); plus S; fixed rows zeroed
 * with 1 on the diagonal.  a: nnz complex; f: n complex. */
int elp_dynamic_matrix(elp_model *m, double omega, double z_re, double z_im,
                       double *a, double *f);

/* element-level operators for a batch of elements of one kind (host in/out):
 * xyz: n x npe x 3; Ke: n x (nd npe)^2; Me: n x npe^2 (the scalar mass
 * pattern; the structural mass is Me (x) I3).  p0..p2 = (rho, c, -) for fluid,
 * (E, nu, rho) for structure.  Fluid matrices are NOT divided by rho here. */
int elp_element_matrices(elp_ctx *ctx, int kind, int64_t n, const double *xyz, double p0,
                         double p1, double p2, double *Ke, double *Me);

/* ---- harmonic sweep ---------------------------------------------------------- */
/* Solves A(w_k) x_k = f_k for every frequency, A and F as elp_dynamic_matrix
 * returns them (MatZeroRows semantics: prescribed values honoured, exactly,
 * on their rows).  X: n_freq x n complex (may be NULL); iters per frequency;
 * relres: FP64 relative residual ||F - A x|| / ||b|| per frequency, where
 * b = F when every prescribed value is 0, and otherwise the right-hand side of
 * the lifted system the solver works on (F - A[:,D] g on the free rows, the
 * values g on the Dirichlet rows D); the iteration stops when the residual is
 * below rtol times the free rows' part of b.  Returns ELP_ENOCONV, with X,
 * relres, iters and stats filled, when any frequency stays above
 * opts->rtol. */
int elp_sweep(elp_model *m, int64_t n_freq, const double *freqs_hz, double z_re,
              double z_im, const elp_solve_opts *opts, double *X, double *relres,
              int32_t *iters, elp_sweep_stats *stats);

/* ---- results at elements and nodes ----------------------------------------- */
/* Sizes of elp_node_results' outputs: nodes, and Brick8/Brick20 elements. */
int elp_results_sizes(const elp_model *m, int64_t *n_nodes, int64_t *n_brick_elements);
/* From solutions X (n_freq x n interleaved complex, as elp_sweep returns
 * them), on the device:
 *  cells (may be null): n_freq x n_brick_elements x 6 complex, the stresses
 *    D B u of each brick at its centroid from its 8 corner nodes, bricks in
 *    ascending element id, components in the order of the strain vector
 *    (xx, yy, zz, xy, yz, xz) as elPaSo stores them in slots 0..5
 *    (cElementStructureBrick::computeStressesCells,
 *    element/structure/linear/brick/elementstructurebrick.cpp:48-158);
 *    the loss factor scales D as in the solve (E(1 + i eta) or E(1 + i w eta))
 *  nodes (may be null): n_freq x n_nodes x 9 complex in ascending node id,
 *    elPaSo's m_StressesNodes layout (eStresses, misc/mytypes.h:79-93):
 *    slots 0..5 the centroid stresses of the node's bricks, each divided by
 *    the number of elements at the node, summed
 *    (elementstructurebrick.cpp:161-279); slots 6..8 grad p of each of the
 *    node's fluid elements evaluated at the node, summed over the elements
 *    (cElementFluid3d::computeStressesNodes / computeGradPPoint,
 *    element/fluid/elementfluid3d.cpp:380-540; elPaSo sums each element's
 *    terms by magnitude, here in local node order: the same to rounding);
 *    other slots 0.
 * Replaces cAnalysis::computeStressesCells / computeStressesNodes
 * (analysis/analysis.cpp:869-937), which run per frequency on the host.
 * kernel_ms (may be null): device time of the kernels (HIP events), without
 * the host copies.
 * ELP_EINVAL for cells of a model mixing Brick8 and Brick20. */
int elp_node_results(elp_model *m, int64_t n_freq, const double *freqs_hz, const double *X,
                     double *cells, double *nodes, double *kernel_ms);

/* Allocation-free variant of elp_node_results, nodes only, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It has no counterpart in the
 * reference.  X_dev: device pointer, n_freq x n interleaved complex
 * solutions; nodes_dev: device pointer, n_freq x n_nodes x 9 complex, the
 * layout of elp_node_results' nodes.  Same kernels and bits as
 * elp_node_results.  Partial results (the brick centroid stresses) go to a
 * static scratch array of 64 MiB per GPU, held for the life of the process:
 * ELP_EINVAL when one frequency of the model's bricks does not fit
 * (about 699,000 brick elements); frequencies are processed in chunks that
 * fit.  The call allocates nothing, runs on the context's stream and
 * returns when the results are written.  Calls on one GPU share its scratch:
 * calls that overlap on one GPU (from several threads) race. */
int elp_node_results_noalloc(elp_model *m, int64_t n_freq, const double *freqs_hz, const double *X_dev,
                             double *nodes_dev);

/* ---- serving: continuous batching of sweep requests ----------------------------- */
int elp_server_create(elp_ctx *ctx, int64_t hbm_budget_bytes, int max_slots, elp_server **out);
void elp_server_destroy(elp_server *s);
/* On success the server takes ownership of m (finalized, assembled or not)
 * and destroys it when the request ends; on failure the caller keeps it.
 * Every elp_server_wait must have returned before elp_server_destroy is
 * called from another thread; requests that have not started when the
 * server is destroyed end with ELP_ESTATE ("server stopped"). */
int elp_server_submit(elp_server *s, elp_model *m, int64_t n_freq, const double *freqs_hz,
                      double z_re, double z_im, const elp_solve_opts *opts, int64_t *job_id);
/* 1 if finished, 0 if pending */
int elp_server_poll(elp_server *s, int64_t job_id);
/* ELP_ENOCONV: the request ran, X / relres / iters are filled, and some
 * frequency stayed above its rtol */
int elp_server_wait(elp_server *s, int64_t job_id, double *X, double *relres, int32_t *iters,
                    double *latency_s);
int elp_server_stats(elp_server *s, int64_t *admitted, int64_t *queued, int64_t *bytes_in_use);

#ifdef __cplusplus
}
#endif
#endif
