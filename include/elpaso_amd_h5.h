/*
 * elpaso_amd_h5.h -- elPaSo HDF5 model input and result output for
 * libelpaso_amd (libelpaso_amd_h5.so, host code only, links libhdf5 1.10).
 *
 * Conventions as in elpaso_amd.h: every function returns 0 or a negative
 * ELP_E* code and never throws; the message of the last failing call of the
 * calling thread is elp_h5_last_error() (errors of the libelpaso_amd calls
 * made while reading a model are copied there too).
 *
 * Reference interfaces replaced (paths relative to elpasoCore/source of
 * elPaSo-Core):
 *  - elp_h5_read_analysis, elp_h5_read_model <- cFemParserHDF5 (misc/parser/
 *    femparserhdf5.cpp:62-560) driven by cFemParserInterface::parseInputFile
 *    (misc/parser/femparserinterface.cpp:189-663), group, dataset and
 *    attribute names of misc/hdf5/rulesh5.h:31-336
 *  - elp_h5_write_solution <- the HDF5 export of cAnalysisFrequencyBasic::FullRun
 *    (analysis/frequency/analysisfrequencybasic.cpp:185-214, 243-265)
 *  - elp_h5_write_csr <- MathLibrary::ExportPetscMatToHDF5Complex
 *    (math/mathlibrary.cpp:53-86)
 */
#ifndef ELPASO_AMD_H5_H
#define ELPASO_AMD_H5_H

#include <stdint.h>

#include "elpaso_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* /Analysis attributes (rulesh5.h:31-44; femparserhdf5.cpp:67-76, 173-194):
 * frequencies of a "frequency-basic" run are f_k = start + delta * k, k < steps
 * (analysisfrequencybasic.cpp:36-72) */
typedef struct elp_h5_analysis {
  char type[64];      /* "type", e.g. "frequency-basic"                     */
  char description[256];
  double start;       /* "start" [Hz]                                       */
  double delta;       /* "delta" [Hz]                                       */
  int64_t steps;      /* "steps"                                            */
  int64_t revision;   /* "revision" (elPaSo reads revision >= 6 only)       */
  int32_t write_hdf5; /* "hdf5": elPaSo writes the solution file when != 0  */
} elp_h5_analysis;

/* What elp_h5_read_model put into the model. */
typedef struct elp_h5_counts {
  int64_t nodes;
  int64_t materials;      /* materials of a kind the library assembles     */
  int64_t element_groups;
  int64_t elements;
  int64_t constraints;    /* /NodeConstraints datasets                     */
  int64_t node_loads;     /* /NodeLoads datasets                           */
  int64_t interfaces;     /* interface elements                            */
} elp_h5_counts;

const char *elp_h5_last_error(void);

/* Reads /Analysis.  ELP_EINVAL for a file that is not an elPaSo input of
 * revision >= 6 (femparserhdf5.cpp:197-211). */
int elp_h5_read_analysis(const char *path, elp_h5_analysis *out);

/* Fills an unfinalized model from an elPaSo input file, in the order of
 * parseInputFile: /Nodes/mtxFemNodes (compound Ids, xCoords, yCoords,
 * zCoords), /Materials (MaterialType AF_LIN_UAF_ISO_DIR: c, rho;
 * STR_LIN_ELA_ISO_DIR: E, nu, rho; STR_LIN_VIS_ISO_DIR: E, nu, rho, eta,
 * type; parameters stored as strings, materialfactory.cpp:41-100), /Elements
 * (datasets of rows [element id, node ids...] with ElementType "Fluid8",
 * "Brick8" -- and this library's "Fluid20", "Brick20" -- and MaterialId),
 * /NodeConstraints ("acoustic": valp on the fluid DOF; "structure": flags
 * u1 ... x2_2 with values valu1 ... valx2_2 on DOFs 0 ... 20,
 * boundaryconditionfactory.cpp:36-60), /NodeLoads ("point_force": Fx, Fy,
 * Fz; "fluid": P as the source strength Q, nodalforcefactory.cpp:36-55) and
 * /InterfaceElements (type "InterfaceMindlin", rows [id, ori, 4 structure
 * nodes, 4 fluid nodes], elementinterfacefactory.cpp:62-98).
 * ELP_EINVAL, naming the item, for anything the file holds that the library
 * does not assemble (another element, material, load or interface type,
 * element loads, non-conforming interfaces) -- nothing is skipped silently.
 * analysis and counts may be null. */
int elp_h5_read_model(elp_model *m, const char *path, elp_h5_analysis *analysis,
                      elp_h5_counts *counts);

/* Writes a solution file as elPaSo's frequency loop does: group
 * /Solution/State with datasets vecFemStep{k+1} (compound {real, imag} of
 * n doubles, row order) and string attributes FreqStep{k+1} =
 * std::to_string(freqs_hz[k]) ("%f"), and /Solution/Maps/mtxDofMap (n x 2
 * int32: node id, DOF enum).  X: n_freq x n interleaved complex, as
 * elp_sweep returns it.  node_ids, dofs: as elp_model_dof_map returns them.
 * The file is created, or truncated when it exists. */
int elp_h5_write_solution(const char *path, int64_t n, int64_t n_freq, const double *freqs_hz,
                          const double *X, const int64_t *node_ids, const int32_t *dofs);

/* Writes /SystemMatrices/<name> as ExportPetscMatToHDF5Complex does:
 * int32 attributes numrow, numcol; int32 datasets vecCsrIa (n + 1 entries)
 * and vecCsrJa (nnz entries), both shifted to 1-based; cmpCsrVal (compound
 * {real, imag}).  rowptr, colind: 0-based CSR as elp_dynamic_matrix gives;
 * vals: nnz interleaved complex.  The file is opened for writing when it
 * exists (other groups are kept; an existing group of this name is an
 * error), else created.  A rowptr that does not start at 0 or decreases, or
 * a column index outside [0, n), is ELP_EINVAL and nothing is written. */
int elp_h5_write_csr(const char *path, const char *name, int64_t n, const int32_t *rowptr,
                     const int32_t *colind, const double *vals);

/* Reads a compound {real, imag} double vector (a vecFemStep or cmpCsrVal
 * dataset) of exactly n entries into X (2n doubles). */
int elp_h5_read_complex_vector(const char *path, const char *dataset, int64_t n, double *X);

#ifdef __cplusplus
}
#endif

#endif
