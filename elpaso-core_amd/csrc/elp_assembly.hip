// Element and global assembly kernels (gfx950, FP64).
//
// Volume elements: one element per wavefront.  The lanes of the wavefront
// first evaluate the Jacobian, its determinant and inverse and the physical
// shape-function gradients at the quadrature points (one lane per point) and
// stage them in LDS; then every lane contracts a strided set of element-matrix
// entries over the quadrature points from LDS.
//
// Global matrices: a deterministic gather, one wavefront per node.  The
// wavefront walks the node's adjacent elements in ascending element id (the
// order in which elPaSo adds element matrices with MatSetValues,
// analysisfrequencybasic.cpp:281-304), adds the node's element-matrix rows into
// LDS accumulators laid out like the node's CSR rows, then writes the rows
// with coalesced stores.  No atomics: repeated runs are bit-identical.
//
// Faces (Robin C_Gamma) and FSI interface elements: one thread per face, then
// a segmented reduction over a precomputed, ordered contribution list.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "elp_fem.h"
#include "elp_internal.h"

namespace elp {

__constant__ HexTables c_tab;

void upload_tables(Ctx *) {
  const HexTables &t = host_tables();
  ELP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &t, sizeof(HexTables)));
}

template <int NPE>
__device__ inline const double *tab_n(int g, int d) {
  if constexpr (NPE == 8) return c_tab.h8[g][d];
  else return c_tab.h20[g][d];
}
template <int NPE>
__device__ inline double tab_w(int g) {
  if constexpr (NPE == 8) return c_tab.w8[g];
  else return c_tab.w27[g];
}

// LDS staging of one element: coordinates and, per quadrature point, w*detJ
// and the physical gradients grad[g][i][k].
template <int NPE, int NG>
struct ElemLds {
  double x[NPE * 3];
  double wdj[NG];
  double grad[NG][3][NPE];
};

// Jacobian, determinant and inverse at one point: J(i,j) = sum_k dN_i[k] x_k[j],
// det by Sarrus' rule, inverse by cofactors (element/elementfem.cpp:158-210);
// gradients dN/dx_i = sum_j invJ(i,j) dN/dxi_j (elementstructurebrick.cpp:481-486).
template <int NPE, int NG>
__device__ inline bool point_geometry(ElemLds<NPE, NG> &L, int g) {
  const double *dxi = tab_n<NPE>(g, 1), *deta = tab_n<NPE>(g, 2), *dzeta = tab_n<NPE>(g, 3);
  double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // partial unrolling keeps the 20-node loops of Hex20 within the register budget
#pragma unroll 4
  for (int k = 0; k < NPE; k++) {
    double x0 = L.x[3 * k], x1 = L.x[3 * k + 1], x2 = L.x[3 * k + 2];
    J[0] += dxi[k] * x0; J[1] += dxi[k] * x1; J[2] += dxi[k] * x2;
    J[3] += deta[k] * x0; J[4] += deta[k] * x1; J[5] += deta[k] * x2;
    J[6] += dzeta[k] * x0; J[7] += dzeta[k] * x1; J[8] += dzeta[k] * x2;
  }
  double det = J[0] * J[4] * J[8] + J[1] * J[5] * J[6] + J[2] * J[3] * J[7] -
               J[2] * J[4] * J[6] - J[0] * J[5] * J[7] - J[1] * J[3] * J[8];
  double iJ[9];
  iJ[0] = (J[4] * J[8] - J[5] * J[7]) / det;
  iJ[3] = -(J[3] * J[8] - J[5] * J[6]) / det;
  iJ[6] = (J[3] * J[7] - J[4] * J[6]) / det;
  iJ[1] = -(J[1] * J[8] - J[2] * J[7]) / det;
  iJ[4] = (J[0] * J[8] - J[2] * J[6]) / det;
  iJ[7] = -(J[0] * J[7] - J[1] * J[6]) / det;
  iJ[2] = (J[1] * J[5] - J[2] * J[4]) / det;
  iJ[5] = -(J[0] * J[5] - J[2] * J[3]) / det;
  iJ[8] = (J[0] * J[4] - J[1] * J[3]) / det;
  L.wdj[g] = tab_w<NPE>(g) * det;
#pragma unroll 4
  for (int k = 0; k < NPE; k++) {
    double a = dxi[k], b = deta[k], c = dzeta[k];
    L.grad[g][0][k] = iJ[0] * a + iJ[1] * b + iJ[2] * c;
    L.grad[g][1][k] = iJ[3] * a + iJ[4] * b + iJ[5] * c;
    L.grad[g][2][k] = iJ[6] * a + iJ[7] * b + iJ[8] * c;
  }
  return det > 0.0;
}

constexpr int kWaves = 4;  // wavefronts (elements / nodes) per workgroup

// Entry t of the upper triangle (a <= b) of a symmetric N x N matrix, N even:
// rows p and N-1-p form one run of N+1 entries, so t decodes with one
// division by a constant.
template <int N>
__device__ inline void tri_entry(int t, int &a, int &b) {
  static_assert(N % 2 == 0, "paired rows need an even order");
  const int p = t / (N + 1), r = t - p * (N + 1);
  if (r < N - p) {
    a = p;
    b = p + r;
  } else {
    a = N - 1 - p;
    b = a + (r - (N - p));
  }
}

// Element matrices of fluid elements: K_ab = sum_g wdJ grad_a.grad_b / rho,
// M_ab = sum_g (1/c^2) N_a N_b wdJ / rho  (elementfluid3d.cpp:54-167, with the
// frequency-independent 1/rho of elPaSo's 1/(rho w^2) scaling, elementfluid.cpp:74-76).
// Structural elements: K = sum_g B^T D B wdJ for isotropic D in closed form
// lambda p_d q_e + mu p_e q_d + mu delta_de p.q, M_ab = rho sum_g N_a N_b wdJ
// (elementstructurebrick.cpp:470-554).  params: fluid (rho, 1/c^2), brick
// (lambda, mu, rho).  bad receives the smallest element index with detJ <= 0.
// K and M are symmetric: each lane computes entries of the upper triangle and
// writes them to both halves.
// ROWS: instead of [element][row][col] blocks, local node a's rows go to the
// record row_at[e][a] of Ke (records of RS = ND NDF + NPE doubles: its ND rows
// of K, then its row of M), so that a node's records are contiguous.
template <int NPE, int ND, int NG, bool ROWS = false>
__global__ __launch_bounds__(64 * kWaves) void element_kernel(
    const double *__restrict__ xyz, const int32_t *__restrict__ conn, const double *__restrict__ params,
    int64_t n_elem, double *__restrict__ Ke, double *__restrict__ Me, int *bad, int xyz_direct,
    const int32_t *__restrict__ row_at) {
  __shared__ ElemLds<NPE, NG> lds[kWaves];
  __shared__ double sN[NG][NPE];  // shape values at the points, shared by the workgroup's elements
  __shared__ int32_t sj[kWaves][NPE];
  for (int i = threadIdx.x; i < NG * NPE; i += blockDim.x) sN[i / NPE][i % NPE] = tab_n<NPE>(i / NPE, 0)[i % NPE];
  __syncthreads();
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t e = (int64_t)blockIdx.x * kWaves + w;
  if (e >= n_elem) return;  // whole wavefront leaves together
  ElemLds<NPE, NG> &L = lds[w];
  for (int t = lane; t < NPE * 3; t += 64) {
    int k = t / 3, d = t % 3;
    L.x[t] = xyz_direct ? xyz[e * NPE * 3 + t] : xyz[(int64_t)conn[e * NPE + k] * 3 + d];
  }
  if constexpr (ROWS) {
    if (lane < NPE) sj[w][lane] = row_at[e * NPE + lane];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < NG) {
    if (!point_geometry<NPE, NG>(L, lane)) atomicMin(bad, (int)e);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double p0 = params[e * 4], p1 = params[e * 4 + 1], p2 = params[e * 4 + 2];
  constexpr int NDF = NPE * ND;
  double *Kout = Ke + e * (int64_t)NDF * NDF;
  double *Mout = Me + e * (int64_t)NPE * NPE;
  if constexpr (ND == 1) {
    for (int t = lane; t < NPE * (NPE + 1) / 2; t += 64) {
      int a, b;
      tri_entry<NPE>(t, a, b);
      double k = 0.0, m = 0.0;
#pragma unroll 3
      for (int g = 0; g < NG; g++) {
        const double wd = L.wdj[g];
        double r = L.grad[g][0][a] * L.grad[g][0][b] + L.grad[g][1][a] * L.grad[g][1][b] +
                   L.grad[g][2][a] * L.grad[g][2][b];
        k += r * wd;
        m += p1 * sN[g][a] * sN[g][b] * wd;
      }
      if constexpr (ROWS) {
        constexpr int RS = NPE + NPE;
        double *Ra = Ke + (int64_t)sj[w][a] * RS, *Rb = Ke + (int64_t)sj[w][b] * RS;
        Ra[b] = Rb[a] = k / p0;
        Ra[NPE + b] = Rb[NPE + a] = m / p0;
      } else {
        Kout[a * NPE + b] = Kout[b * NPE + a] = k / p0;
        Mout[a * NPE + b] = Mout[b * NPE + a] = m / p0;
      }
    }
  } else {
    const double lam = p0, mu = p1, rho = p2;
    for (int t = lane; t < NDF * (NDF + 1) / 2; t += 64) {
      int row, col;
      tri_entry<NDF>(t, row, col);
      const int a = row / 3, d = row % 3, b = col / 3, dd = col % 3;
      double k = 0.0;
#pragma unroll 3
      for (int g = 0; g < NG; g++) {
        const double pd = L.grad[g][d][a], qe = L.grad[g][dd][b];
        const double pe = L.grad[g][dd][a], qd = L.grad[g][d][b];
        double v = lam * pd * qe + mu * pe * qd;
        if (d == dd)
          v += mu * (L.grad[g][0][a] * L.grad[g][0][b] + L.grad[g][1][a] * L.grad[g][1][b] +
                     L.grad[g][2][a] * L.grad[g][2][b]);
        k += v * L.wdj[g];
      }
      if constexpr (ROWS) {
        constexpr int RS = ND * NDF + NPE;
        Ke[(int64_t)sj[w][a] * RS + d * NDF + col] = k;
        Ke[(int64_t)sj[w][b] * RS + dd * NDF + row] = k;
      } else {
        Kout[row * NDF + col] = Kout[col * NDF + row] = k;
      }
    }
    for (int t = lane; t < NPE * (NPE + 1) / 2; t += 64) {
      int a, b;
      tri_entry<NPE>(t, a, b);
      double m = 0.0;
#pragma unroll 3
      for (int g = 0; g < NG; g++) m += rho * sN[g][a] * sN[g][b] * L.wdj[g];
      if constexpr (ROWS) {
        constexpr int RS = ND * NDF + NPE;
        Ke[(int64_t)sj[w][a] * RS + ND * NDF + b] = Ke[(int64_t)sj[w][b] * RS + ND * NDF + a] = m;
      } else {
        Mout[a * NPE + b] = Mout[b * NPE + a] = m;
      }
    }
  }
}

// Fluid20 element matrices of the assembly on the FP64 matrix cores, in the
// node-ordered records of element_kernel's ROWS layout: K = G diag(w) G^T
// and M = N diag(p1 w) N^T, G (20 x 81) the gradient components at the 27
// points (column 3 g + i), N (20 x 27) the shape values, w = w_g detJ_g.
// Only the 27 inverse Jacobians are kept per element (2.7 KB of LDS per
// wavefront, the shape tables are shared by the workgroup); each MFMA
// operand G[a][3 g + i] = sum_j invJ_g(i, j) dN_a/dxi_j is formed in the
// loop.  v_mfma_f64_16x16x4f64 tiles: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives D[l / 16 + 4 r][l % 16] in element r.
// Tiles (0,0), (0,1), (1,1) of the padded 32 x 32 result; each entry a <= b
// comes from one product and is written to both halves (exactly symmetric).
struct Fluid20Lds {
  double x[20 * 3];
  double wdj[27];
  double ij[27][9];
  int32_t rec[20];
};

__global__ __launch_bounds__(64 * kWaves) void fluid20_mfma_kernel(
    const double *__restrict__ xyz, const int32_t *__restrict__ conn, const double *__restrict__ params,
    int64_t n_elem, double *__restrict__ R, int *bad, const int32_t *__restrict__ row_at) {
  constexpr int NPE = 20, NG = 27, RS = 2 * NPE;
  __shared__ Fluid20Lds lds[kWaves];
  __shared__ double sN[NG][NPE];
  __shared__ double sD[NG][3][NPE];
  for (int i = threadIdx.x; i < NG * NPE; i += blockDim.x) sN[i / NPE][i % NPE] = tab_n<NPE>(i / NPE, 0)[i % NPE];
  for (int i = threadIdx.x; i < NG * 3 * NPE; i += blockDim.x) {
    const int g = i / (3 * NPE), d = (i / NPE) % 3, k = i % NPE;
    sD[g][d][k] = tab_n<NPE>(g, d + 1)[k];
  }
  __syncthreads();
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t e = (int64_t)blockIdx.x * kWaves + w;
  if (e >= n_elem) return;  // whole wavefront leaves together
  Fluid20Lds &L = lds[w];
  for (int t = lane; t < NPE * 3; t += 64) L.x[t] = xyz[(int64_t)conn[e * NPE + t / 3] * 3 + t % 3];
  if (lane < NPE) L.rec[lane] = row_at[e * NPE + lane];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < NG) {
    // Jacobian, Sarrus determinant, cofactor inverse as point_geometry
    const int g = lane;
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (int k = 0; k < NPE; k++) {
      const double x0 = L.x[3 * k], x1 = L.x[3 * k + 1], x2 = L.x[3 * k + 2];
      const double a = sD[g][0][k], b = sD[g][1][k], c = sD[g][2][k];
      J[0] += a * x0; J[1] += a * x1; J[2] += a * x2;
      J[3] += b * x0; J[4] += b * x1; J[5] += b * x2;
      J[6] += c * x0; J[7] += c * x1; J[8] += c * x2;
    }
    const double det = J[0] * J[4] * J[8] + J[1] * J[5] * J[6] + J[2] * J[3] * J[7] -
                       J[2] * J[4] * J[6] - J[0] * J[5] * J[7] - J[1] * J[3] * J[8];
    double *iJ = L.ij[g];
    iJ[0] = (J[4] * J[8] - J[5] * J[7]) / det;
    iJ[3] = -(J[3] * J[8] - J[5] * J[6]) / det;
    iJ[6] = (J[3] * J[7] - J[4] * J[6]) / det;
    iJ[1] = -(J[1] * J[8] - J[2] * J[7]) / det;
    iJ[4] = (J[0] * J[8] - J[2] * J[6]) / det;
    iJ[7] = -(J[0] * J[7] - J[1] * J[6]) / det;
    iJ[2] = (J[1] * J[5] - J[2] * J[4]) / det;
    iJ[5] = -(J[0] * J[5] - J[2] * J[3]) / det;
    iJ[8] = (J[0] * J[4] - J[1] * J[3]) / det;
    L.wdj[g] = tab_w<NPE>(g) * det;
    if (!(det > 0.0)) atomicMin(bad, (int)e);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double p0 = params[e * 4], p1 = params[e * 4 + 1];
  typedef double v4d __attribute__((ext_vector_type(4)));
  const int li = lane % 16, lk = lane / 16;
  v4d k00 = {0, 0, 0, 0}, k01 = {0, 0, 0, 0}, k11 = {0, 0, 0, 0};
  v4d m00 = {0, 0, 0, 0}, m01 = {0, 0, 0, 0}, m11 = {0, 0, 0, 0};
#pragma unroll 3
  for (int st = 0; st < 21; st++) {
    const int kk = 4 * st + lk;
    const bool kv = kk < 3 * NG;
    const int g = kv ? kk / 3 : 0, i = kk % 3;
    const double *q = &L.ij[g][3 * i];
    const double wk = kv ? L.wdj[g] : 0.0;
    const double a0 = kv ? q[0] * sD[g][0][li] + q[1] * sD[g][1][li] + q[2] * sD[g][2][li] : 0.0;
    const int l1 = li < 4 ? 16 + li : li;  // in-range index; the value is zeroed below
    const double a1v = q[0] * sD[g][0][l1] + q[1] * sD[g][1][l1] + q[2] * sD[g][2][l1];
    const double a1 = (kv && li < 4) ? a1v : 0.0;
    k00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, wk * a0, k00, 0, 0, 0);
    k01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, wk * a1, k01, 0, 0, 0);
    k11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, wk * a1, k11, 0, 0, 0);
  }
#pragma unroll
  for (int st = 0; st < 7; st++) {
    const int g = 4 * st + lk;
    const bool gv = g < NG;
    const int gc = gv ? g : 0;
    const double wk = gv ? p1 * L.wdj[gc] : 0.0;
    const int l1 = li < 4 ? 16 + li : li;  // in-range index; the value is zeroed below
    const double a0 = gv ? sN[gc][li] : 0.0;
    const double a1 = (gv && li < 4) ? sN[gc][l1] : 0.0;
    m00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, wk * a0, m00, 0, 0, 0);
    m01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, wk * a1, m01, 0, 0, 0);
    m11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, wk * a1, m11, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int io = lk + 4 * r;  // row within the tile
    auto put = [&](int a, int b, double kv, double mv) {
      double *Ra = R + (int64_t)L.rec[a] * RS, *Rb = R + (int64_t)L.rec[b] * RS;
      Ra[b] = Rb[a] = kv;
      Ra[NPE + b] = Rb[NPE + a] = mv;
    };
    if (io <= li) put(io, li, k00[r] / p0, m00[r] / p0);
    if (li < 4) put(io, 16 + li, k01[r] / p0, m01[r] / p0);
    if (io < 4 && li < 4 && io <= li) put(16 + io, 16 + li, k11[r] / p0, m11[r] / p0);
  }
}

// Deterministic gather of the element rows into the node's CSR rows (K, M
// and optionally C = alpha K + beta M), one wavefront per node of a 20-node
// kind.  The node's records (element_kernel's / fluid20_mfma_kernel's ROWS
// layout) are contiguous, in ascending element id, with the neighbour slot of
// each record column precomputed (adj_slot).  Each lane owns neighbour slots
// lane, lane + 64, ...: per batch of up to 8 records the wavefront builds the
// inverse map (record, slot) -> record column in LDS, then every lane adds
// its slots' values record by record, in ascending element id, in registers,
// and finally writes its slots' columns of the node's rows.
constexpr int kMaxCols = 256;
constexpr int kGatherBatch = 8;
template <int NPE, int ND>
__global__ __launch_bounds__(64 * kWaves) void gather_kernel(
    int32_t n_nodes, const int32_t *__restrict__ nodes, const int32_t *__restrict__ adj_ptr,
    const uint8_t *__restrict__ adj_slot, const double *__restrict__ R, const int32_t *__restrict__ nbr_ptr,
    const int32_t *__restrict__ nbr_off, const int32_t *__restrict__ node_row, const int32_t *__restrict__ rowptr,
    double *__restrict__ Kv, double *__restrict__ Mv, double *__restrict__ Cv, double alpha, double beta) {
  constexpr int NDF = NPE * ND, RS = ND * NDF + NPE, SP = kMaxCols / 64;
  __shared__ uint8_t inv[kWaves][kGatherBatch][kMaxCols];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int32_t s_node = blockIdx.x * kWaves + w;
  if (s_node >= n_nodes) return;
  const int32_t i = nodes[s_node];
  const int32_t nb0 = nbr_ptr[i], deg = nbr_ptr[i + 1] - nb0;
  const int32_t r0 = node_row[i];
  const int32_t len = rowptr[r0 + 1] - rowptr[r0];
  const int32_t pa0 = adj_ptr[s_node], pa1 = adj_ptr[s_node + 1];
  double ak[SP][ND][ND], am[SP];
#pragma unroll
  for (int p = 0; p < SP; p++) {
    am[p] = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++)
#pragma unroll
      for (int dd = 0; dd < ND; dd++) ak[p][d][dd] = 0.0;
  }
  for (int32_t pb = pa0; pb < pa1; pb += kGatherBatch) {
    const int nk = min(kGatherBatch, pa1 - pb);
    for (int t = lane; t < nk * kMaxCols; t += 64) inv[w][t / kMaxCols][t % kMaxCols] = 0xff;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // the slots of one record are distinct
    for (int it = lane; it < nk * NPE; it += 64) inv[w][it / NPE][adj_slot[(int64_t)pb * NPE + it]] = (uint8_t)(it % NPE);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int kk = 0; kk < nk; kk++) {
      const double *rec = R + (int64_t)(pb + kk) * RS;
#pragma unroll
      for (int p = 0; p < SP; p++) {
        const int sl = lane + 64 * p;
        if (sl >= deg) continue;
        const int b = inv[w][kk][sl];
        if (b == 0xff) continue;
#pragma unroll
        for (int d = 0; d < ND; d++)
#pragma unroll
          for (int dd = 0; dd < ND; dd++) ak[p][d][dd] += rec[d * NDF + b * ND + dd];
        am[p] += rec[ND * NDF + b];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
#pragma unroll
  for (int p = 0; p < SP; p++) {
    const int sl = lane + 64 * p;
    if (sl >= deg) continue;
    const int32_t o0 = nbr_off[nb0 + sl], o1 = sl + 1 < deg ? nbr_off[nb0 + sl + 1] : len;
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const int64_t base = rowptr[r0 + d];
      for (int32_t q = o0; q < o1; q++) {
        const int dd = q - o0;  // DOF within the neighbour
        double kv = 0.0;
#pragma unroll
        for (int e2 = 0; e2 < ND; e2++)
          if (dd == e2) kv = ak[p][d][e2];
        const double mv = (dd == d) ? am[p] : 0.0;
        Kv[base + q] = kv;
        Mv[base + q] = mv;
        if (Cv) Cv[base + q] = alpha * kv + beta * mv;
      }
    }
  }
}

// Robin face matrices C_f(z,s) = sum_g N_z N_s w detJ_face,
// detJ_face = sqrt(|a1|^2 |a2|^2 - (a1.a2)^2) (elementfluid3d.cpp:169-234).
template <int NFN, int NGF>
__global__ void face_kernel(int64_t n_faces, const int32_t *__restrict__ fnodes,
                            const double *__restrict__ xyz, double *__restrict__ out, int *bad) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  double x[NFN][3];
  for (int k = 0; k < NFN; k++)
    for (int d = 0; d < 3; d++) x[k][d] = xyz[(int64_t)fnodes[f * NFN + k] * 3 + d];
  double C[NFN * NFN];
  for (int t = 0; t < NFN * NFN; t++) C[t] = 0.0;
  for (int g = 0; g < NGF; g++) {
    const double *N, *Nx, *Ne;
    double w;
    if constexpr (NFN == 4) {
      N = c_tab.q4[g][0]; Nx = c_tab.q4[g][1]; Ne = c_tab.q4[g][2]; w = c_tab.wq4[g];
    } else {
      N = c_tab.q8[g][0]; Nx = c_tab.q8[g][1]; Ne = c_tab.q8[g][2]; w = c_tab.wq9[g];
    }
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    for (int k = 0; k < NFN; k++)
      for (int d = 0; d < 3; d++) {
        a1[d] += Nx[k] * x[k][d];
        a2[d] += Ne[k] * x[k][d];
      }
    const double a11 = a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2];
    const double a22 = a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2];
    const double a12 = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
    const double det = sqrt(a11 * a22 - a12 * a12);
    if (!(det > 0.0)) atomicMin(bad, (int)f);
    for (int z = 0; z < NFN; z++)
      for (int s = 0; s < NFN; s++) C[z * NFN + s] += N[z] * N[s] * w * det;
  }
  for (int t = 0; t < NFN * NFN; t++) out[f * NFN * NFN + t] = C[t];
}

// FSI coupling of one Quad4 interface element: C(3z+2, s) = sum_g N_z N_s w
// detJ, then projected on the nodal normals a1 x a2 / |a1 x a2| and scaled by
// -ori (elementinterfacemindlin.cpp:80-181, elementinterfacestructural.cpp:209-237).
__global__ void fsi_kernel(int64_t n_if, const int32_t *__restrict__ fnodes,
                           const double *__restrict__ xyz, const double *__restrict__ factor,
                           double *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_if) return;
  double x[4][3];
  for (int k = 0; k < 4; k++)
    for (int d = 0; d < 3; d++) x[k][d] = xyz[(int64_t)fnodes[e * 4 + k] * 3 + d];
  double Cz[4][4];
  for (int z = 0; z < 4; z++)
    for (int s = 0; s < 4; s++) Cz[z][s] = 0.0;
  for (int g = 0; g < 4; g++) {
    const double *N = c_tab.q4[g][0], *Nx = c_tab.q4[g][1], *Ne = c_tab.q4[g][2];
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) {
        a1[d] += Nx[k] * x[k][d];
        a2[d] += Ne[k] * x[k][d];
      }
    const double a11 = a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2];
    const double a22 = a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2];
    const double a12 = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
    const double det = sqrt(a11 * a22 - a12 * a12);
    for (int z = 0; z < 4; z++)
      for (int s = 0; s < 4; s++) Cz[z][s] += N[z] * N[s] * c_tab.wq4[g] * det;
  }
  const double fac = factor[e];
  for (int n = 0; n < 4; n++) {
    const double *Nx = c_tab.q4n[n][0], *Ne = c_tab.q4n[n][1];
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) {
        a1[d] += Nx[k] * x[k][d];
        a2[d] += Ne[k] * x[k][d];
      }
    double nv[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2],
                    a1[0] * a2[1] - a1[1] * a2[0]};
    const double inv = 1.0 / sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    for (int d = 0; d < 3; d++) nv[d] *= inv;
    for (int s = 0; s < 4; s++)
      for (int d = 0; d < 3; d++) out[e * 48 + (3 * n + d) * 4 + s] = nv[d] * Cz[n][s] * fac;
  }
}

// out[k] = sum of buf over the ordered contribution list of entry k
__global__ void segmented_sum_kernel(int64_t n, const int32_t *__restrict__ ptr,
                                     const int32_t *__restrict__ idx, const double *__restrict__ buf,
                                     double *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s = 0.0;
  for (int32_t t = ptr[k]; t < ptr[k + 1]; t++) s += buf[idx[t]];
  out[k] = s;
}

static int blocks_for(int64_t n, int per) { return (int)((n + per - 1) / per); }

// rows: write the model's node-ordered records (kb.d_row_at) instead of
// [element][row][col] blocks (assembly of the 20-node kinds)
static void launch_element(const KindBlock &kb, const double *xyz, int xyz_direct, double *Ke,
                           double *Me, int *bad, hipStream_t s, bool rows = false) {
  const int grid = blocks_for(kb.n_elem, kWaves);
  if (grid == 0) return;
  const int32_t *conn = kb.d_conn.p;
  const double *par = kb.d_params.p;
  const int32_t *ra = kb.d_row_at.p;
  switch (kb.kind) {
    case ELP_FLUID8:
      hipLaunchKernelGGL((element_kernel<8, 1, 8>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                         kb.n_elem, Ke, Me, bad, xyz_direct, nullptr);
      break;
    case ELP_BRICK8:
      hipLaunchKernelGGL((element_kernel<8, 3, 8>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                         kb.n_elem, Ke, Me, bad, xyz_direct, nullptr);
      break;
    case ELP_FLUID20:
      if (rows)
        hipLaunchKernelGGL(fluid20_mfma_kernel, dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par, kb.n_elem, Ke,
                           bad, ra);
      else
        hipLaunchKernelGGL((element_kernel<20, 1, 27>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                           kb.n_elem, Ke, Me, bad, xyz_direct, nullptr);
      break;
    default:
      if (rows)
        hipLaunchKernelGGL((element_kernel<20, 3, 27, true>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                           kb.n_elem, Ke, Me, bad, xyz_direct, ra);
      else
        hipLaunchKernelGGL((element_kernel<20, 3, 27>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                           kb.n_elem, Ke, Me, bad, xyz_direct, nullptr);
      break;
  }
  ELP_HIP(hipGetLastError());
}

static void launch_gather(Model &m, KindBlock &kb, hipStream_t s) {
  const int32_t nn = (int32_t)kb.nodes.size();
  const int grid = blocks_for(nn, kWaves);
  if (grid == 0) return;
  double *Cv = m.has_rayleigh ? m.d_C.p : nullptr;
#define ELP_GATHER_ARGS                                                                                          \
  nn, kb.d_nodes.p, kb.d_adj_ptr.p, kb.d_adj_slot.p, kb.d_Ke.p, m.d_nbr_ptr.p, m.d_nbr_off.p, m.d_node_row.p, \
      m.d_rowptr.p, m.d_K.p, m.d_M.p, Cv, m.ray_a, m.ray_b
  if (kb.kind == ELP_FLUID20)
    hipLaunchKernelGGL((gather_kernel<20, 1>), dim3(grid), dim3(64 * kWaves), 0, s, ELP_GATHER_ARGS);
  else
    hipLaunchKernelGGL((gather_kernel<20, 3>), dim3(grid), dim3(64 * kWaves), 0, s, ELP_GATHER_ARGS);
#undef ELP_GATHER_ARGS
  ELP_HIP(hipGetLastError());
}

// Hex8 trilinear factors at the 2x2x2 Gauss points.  Node q's natural
// coordinate signs are bits of 0x66 (x), 0xCC (y), 0xF0 (z); point g has the
// indices g % 2, (g / 2) % 2, g / 4 (index 0 = +0.577350269189626).  The 1-D
// factor of a node at a point is 1 + 0.577350269189626 where the node's sign
// bit differs from the point's index and 1 - 0.577350269189626 where it is
// equal.  Derivatives +-(f_y f_z) * 0.125 (and cyclic) are the arithmetic of
// the host table c_tab.h8, so the Jacobian is bit-identical to a table-driven
// one.
constexpr double kH8Hi = 1. + 0.577350269189626, kH8Lo = 1. - 0.577350269189626;
__host__ __device__ constexpr int h8_sx(int q) { return (0x66 >> q) & 1; }
__host__ __device__ constexpr int h8_sy(int q) { return (0xCC >> q) & 1; }
__host__ __device__ constexpr int h8_sz(int q) { return (0xF0 >> q) & 1; }
__device__ inline double h8_f(int s, int ix) { return s != ix ? kH8Hi : kH8Lo; }

// Fused assembly of trilinear (Hex8) elements directly into the CSR rows of
// the nodes, without element matrices in HBM.  A persistent wavefront takes
// its nodes one after another; the node's adjacent elements are taken in
// batches of 8 in ascending element id, and per batch:
//  0. lane (k, q) stages the coordinates of node q of element k in LDS;
//  1. lane (k, g) computes the geometry of element k at quadrature point g
//     (Jacobian, determinant, inverse) and the physical gradients of the 8
//     shape functions there, staged in LDS with w detJ;
//  2. lane (k, b) contracts element k's block of the (node a, neighbour b)
//     pair over the 8 points: P = sum_g w detJ grad N_a (x) grad N_b gives
//     the block lambda P + mu P^T + mu tr(P) I (Brick8) or tr(P) / rho
//     (Fluid8), plus the mass term sum_g w detJ N_a N_b;
//  3. the blocks go to LDS and the lanes owning a neighbour slot (two lanes
//     per slot for 3 DOF, one per half of the block) add them into registers
//     in ascending element id, the order of elPaSo's MatSetValues.
// After the last batch the node's K, M (and C) entries go to its CSR rows:
// on the common path through an LDS image of the rows and coalesced stores,
// on the general path straight from the slot lanes.  No atomics on values:
// repeated runs are bit-identical.
//
// The inputs of a node sit four dependent loads deep (descriptor ->
// adjacency -> connectivity -> coordinates).  For the nodes in at most 8
// elements with at most one pass of slots (32 for 3 DOF, 64 for 1) and no
// neighbour of another field (all nodes of a structured mesh away from
// interfaces) the wavefront issues those of its next two nodes between the
// phases of the current one, so their latency hides behind its arithmetic,
// and runs one batch and one pass.  The others (GENERAL) load their inputs
// in place and run the batches once per group of slots.
template <int ND>
struct Hex8Lds {
  // lane (k, b)'s block at an odd stride of doubles (no LDS bank conflicts):
  // 3 DOF: half 0 = K00 K01 K02 K10 K11 M, half 1 = K12 K20 K21 K22 0 M;
  // 1 DOF: K, M
  static constexpr int kRec = ND == 1 ? 3 : 13, kHalf = ND == 1 ? 2 : 6;
  static constexpr int kRow = 25;  // per (element k, point g): 3 x 8 derivatives and w detJ
  union {
    double xe[3][64];       // phase 0-1: coordinate c of node q of element k, [c][k*8+q]
    double gr[64 * kRow];   // phase 1-2: derivative c of N_q at point g of element k at
                            // [(k*8+g)*25 + c*8 + q], w detJ at [(k*8+g)*25 + 24]; the odd
                            // row stride keeps the lanes' rows in distinct LDS banks
    double v[64][kRec];     // phase 3: the block of lane (k, b)
    double rows[ND * 96 + 32];  // phase 4: image of the node's rows and the mass per slot
  } u;
  uint64_t slots[64];  // slot sp of the pass: byte k = b of batch element k touching it, 0xFF if none
  __device__ double *point(int k, int g) { return u.gr + (k * 8 + g) * kRow; }
  __device__ uint64_t &cb(int sp) { return slots[sp]; }
};

// The chain of one node's inputs that is loaded ahead (lane (k, j) holds
// element k's share); what the node needs only from its phase 2 on (local
// index, slots, material, row offsets) is loaded when its work starts.
struct Hex8Fetch {
  int32_t raw;                                // descriptor field lane & 7
  int32_t a0, nadj, nb0, deg, rowbase, len;  // descriptor (wave-uniform)
  int32_t e, nq;                              // element k, rank of its node j
  double x0, x1, x2;                          // coordinates of node j
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int ND, bool GENERAL, bool WITH_C>
__global__ __launch_bounds__(64) void hex8_node_kernel(
    int32_t n_nodes, const int32_t *__restrict__ desc, const int32_t *__restrict__ adj_elem,
    const uint8_t *__restrict__ adj_local, const uint8_t *__restrict__ adj_slot, const int32_t *__restrict__ conn,
    const double *__restrict__ params, const double *__restrict__ xyz, const int32_t *__restrict__ nbr_off,
    double *__restrict__ Kv, double *__restrict__ Mv, double *__restrict__ Cv, double alpha, double beta,
    int *bad) {
  constexpr int H = ND == 1 ? 1 : 2;  // lanes per neighbour slot
  constexpr int SPP = 64 / H;         // slots per pass
  constexpr int RH = Hex8Lds<ND>::kHalf;
  __shared__ Hex8Lds<ND> L;
  const int lane = threadIdx.x;
  const int k = lane >> 3, j = lane & 7;  // j: node q (phase 0), point g (1), neighbour b (2)
  const int gxi = j & 1, gyi = (j >> 1) & 1, gzi = j >> 2;
  int32_t bad_elem = INT32_MAX;  // lowest element with detJ <= 0 seen by this lane

  // the nodes are cut into 8 ranges, one per XCD (the hardware deals
  // workgroups round-robin to the XCDs): neighbouring nodes, which share
  // elements and coordinates, run on one XCD at the same time
  const int32_t wx = (int32_t)(gridDim.x >> 3);
  const int32_t per = (n_nodes + 7) >> 3;
  const int32_t beg = (int32_t)(blockIdx.x & 7) * per, end = min(n_nodes, beg + per);
  int32_t t = beg + (int32_t)(blockIdx.x >> 3);
  if (t >= end) return;

  // the four load levels of node tt
  auto fetch1 = [&](Hex8Fetch &f, int32_t tt) { f.raw = desc[(int64_t)tt * 8 + j]; };
  auto fetch2 = [&](Hex8Fetch &f) {
    f.a0 = __builtin_amdgcn_readlane(f.raw, 0);
    f.nadj = __builtin_amdgcn_readlane(f.raw, 1);
    f.nb0 = __builtin_amdgcn_readlane(f.raw, 2);
    f.deg = __builtin_amdgcn_readlane(f.raw, 3);
    f.rowbase = __builtin_amdgcn_readlane(f.raw, 4);
    f.len = __builtin_amdgcn_readlane(f.raw, 5);
    // Every lane loads, lanes past the node's elements repeating the last
    // one: loads under a branch would make the number of loads in flight
    // path-dependent, and the compiler then waits for all of them wherever
    // an earlier-loaded value is used.
    f.e = adj_elem[f.a0 + min(k, f.nadj - 1)];
  };
  auto fetch3 = [&](Hex8Fetch &f) { f.nq = conn[(int64_t)f.e * 8 + j]; };
  auto fetch4 = [&](Hex8Fetch &f) {
    f.x0 = xyz[(int64_t)f.nq * 3];
    f.x1 = xyz[(int64_t)f.nq * 3 + 1];
    f.x2 = xyz[(int64_t)f.nq * 3 + 2];
  };

  // node c's work; the hooks issue the loads of the next nodes between its
  // phases (at the start, after the geometry, after the contraction)
  auto process = [&](const Hex8Fetch &c, auto &&at_start, auto &&after_geometry, auto &&after_blocks) {
    const int npass = GENERAL ? (c.deg + SPP - 1) / SPP : 1;
    const int nbatch = GENERAL ? (c.nadj + 7) / 8 : 1;
    for (int ip = 0; ip < npass; ip++) {
      const int s0 = ip * SPP;
      const int s = s0 + lane / H, h = lane % H;  // slot of this lane, half of its block
      // offsets of the lane's slot and of the next one in the rows (the width
      // is taken where the rows are written: using loaded values early would
      // wait for them)
      const int32_t off = GENERAL ? nbr_off[c.nb0 + min(s, c.deg - 1)] : 0;
      const int32_t offn = GENERAL ? nbr_off[c.nb0 + min(s + 1, c.deg - 1)] : 0;
      double acc[RH];
#pragma unroll
      for (int u = 0; u < RH; u++) acc[u] = 0.0;
      for (int ib = 0; ib < nbatch; ib++) {
        const int base = ib * 8;
        const int nk = min(8, c.nadj - base);
        const bool first = s0 == 0 && base == 0;  // the pass and batch of the prefetched inputs
        // batch inputs: the coordinate chain prefetched for the first batch,
        // loaded here for the others; the rest is used from phase 2 on
        const int32_t p = c.a0 + base + min(k, nk - 1);
        int32_t e = c.e;
        double x0 = c.x0, x1 = c.x1, x2 = c.x2;
        if (GENERAL && !first) {
          e = adj_elem[p];
          const int32_t nq = conn[(int64_t)e * 8 + j];
          x0 = xyz[(int64_t)nq * 3];
          x1 = xyz[(int64_t)nq * 3 + 1];
          x2 = xyz[(int64_t)nq * 3 + 2];
        }
        const int32_t a = adj_local[p], slot = adj_slot[(int64_t)p * 8 + j];
        const double pl = params[(int64_t)e * 4], pm = params[(int64_t)e * 4 + 1], pr = params[(int64_t)e * 4 + 2];
        wave_sync();  // the previous batch's blocks are read
        if (first) at_start();
        if (lane < SPP) L.cb(lane) = ~0ull;
        if (k < nk) {
          L.u.xe[0][lane] = x0;
          L.u.xe[1][lane] = x1;
          L.u.xe[2][lane] = x2;
        }
        wave_sync();
        // 1. geometry of (element k, point g = j): J(i,j) = sum_q dN_q/dxi_i x_q[j],
        //    det by Sarrus' rule, inverse by cofactors, gradients invJ dN
        //    (element/elementfem.cpp:158-210); derivative magnitudes
        //    yz[sy][sz] = (f_y f_z) / 8 etc. of lane j as point g
        double yz[2][2], xz[2][2], xy[2][2];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
          for (int v = 0; v < 2; v++) {
            yz[u][v] = (h8_f(u, gyi) * h8_f(v, gzi)) * 0.125;
            xz[u][v] = (h8_f(u, gxi) * h8_f(v, gzi)) * 0.125;
            xy[u][v] = (h8_f(u, gxi) * h8_f(v, gyi)) * 0.125;
          }
        double G[9], wdj = 0.0;
        if (k < nk) {
          double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const double c0 = L.u.xe[0][k * 8 + q], c1 = L.u.xe[1][k * 8 + q], c2 = L.u.xe[2][k * 8 + q];
            const double dx = (h8_sx(q) ? 1.0 : -1.0) * yz[h8_sy(q)][h8_sz(q)];
            const double dy = (h8_sy(q) ? 1.0 : -1.0) * xz[h8_sx(q)][h8_sz(q)];
            const double dz = (h8_sz(q) ? 1.0 : -1.0) * xy[h8_sx(q)][h8_sy(q)];
            J[0] += dx * c0; J[1] += dx * c1; J[2] += dx * c2;
            J[3] += dy * c0; J[4] += dy * c1; J[5] += dy * c2;
            J[6] += dz * c0; J[7] += dz * c1; J[8] += dz * c2;
          }
          const double det = J[0] * J[4] * J[8] + J[1] * J[5] * J[6] + J[2] * J[3] * J[7] -
                             J[2] * J[4] * J[6] - J[0] * J[5] * J[7] - J[1] * J[3] * J[8];
          if (!(det > 0.0)) bad_elem = min(bad_elem, e);
          const double id = 1.0 / det;
          G[0] = (J[4] * J[8] - J[5] * J[7]) * id;
          G[3] = -(J[3] * J[8] - J[5] * J[6]) * id;
          G[6] = (J[3] * J[7] - J[4] * J[6]) * id;
          G[1] = -(J[1] * J[8] - J[2] * J[7]) * id;
          G[4] = (J[0] * J[8] - J[2] * J[6]) * id;
          G[7] = -(J[0] * J[7] - J[1] * J[6]) * id;
          G[2] = (J[1] * J[5] - J[2] * J[4]) * id;
          G[5] = -(J[0] * J[5] - J[2] * J[3]) * id;
          G[8] = (J[0] * J[4] - J[1] * J[3]) * id;
          wdj = det;  // w = 1 at every point of the 2x2x2 rule
        }
        wave_sync();  // coordinates read: the gradients overwrite them
        if (k < nk) {
          double *row = L.point(k, j);
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const double dx = (h8_sx(q) ? 1.0 : -1.0) * yz[h8_sy(q)][h8_sz(q)];
            const double dy = (h8_sy(q) ? 1.0 : -1.0) * xz[h8_sx(q)][h8_sz(q)];
            const double dz = (h8_sz(q) ? 1.0 : -1.0) * xy[h8_sx(q)][h8_sy(q)];
            row[q] = G[0] * dx + G[1] * dy + G[2] * dz;
            row[8 + q] = G[3] * dx + G[4] * dy + G[5] * dz;
            row[16 + q] = G[6] * dx + G[7] * dy + G[8] * dz;
          }
          row[24] = wdj;
        }
        if (first) after_geometry();
        wave_sync();
        // 2. unit (k, b = j): element k's block of the (node, neighbour b) pair
        double V[Hex8Lds<ND>::kRec];
        if (k < nk) {
          // N_a N_b / 64 at point (ix, iy, iz) = pxy[ix + 2 iy] * pz[iz]
          const double px0 = h8_f(h8_sx(a), 0) * h8_f(h8_sx(j), 0), px1 = h8_f(h8_sx(a), 1) * h8_f(h8_sx(j), 1);
          const double py0 = h8_f(h8_sy(a), 0) * h8_f(h8_sy(j), 0), py1 = h8_f(h8_sy(a), 1) * h8_f(h8_sy(j), 1);
          const double pz[2] = {h8_f(h8_sz(a), 0) * h8_f(h8_sz(j), 0) * 0.015625,
                                h8_f(h8_sz(a), 1) * h8_f(h8_sz(j), 1) * 0.015625};
          const double pxy[4] = {px0 * py0, px1 * py0, px0 * py1, px1 * py1};
          double P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, mm = 0.0, kk = 0.0;
#pragma unroll
          for (int gg = 0; gg < 8; gg++) {
            const double *row = L.point(k, gg);
            const double ax = row[a], ay = row[8 + a], az = row[16 + a];
            const double bx = row[j], by = row[8 + j], bz = row[16 + j];
            const double wd = row[24];
            const double w0 = wd * ax, w1 = wd * ay, w2 = wd * az;
            mm += (pxy[gg & 3] * pz[gg >> 2]) * wd;
            if constexpr (ND == 1) {
              kk += w0 * bx + w1 * by + w2 * bz;
            } else {
              P[0] += w0 * bx; P[1] += w0 * by; P[2] += w0 * bz;
              P[3] += w1 * bx; P[4] += w1 * by; P[5] += w1 * bz;
              P[6] += w2 * bx; P[7] += w2 * by; P[8] += w2 * bz;
            }
            // one point's LDS reads at a time: hoisting all 8 would cost
            // registers, and occupancy, for latency the other waves cover
            __builtin_amdgcn_sched_barrier(0);
          }
          // the element's own material law (params: Fluid8 rho, 1/c^2; Brick8
          // lambda, mu, rho)
          if constexpr (ND == 1) {
            V[0] = kk / pl;
            V[1] = pm * mm / pl;
            V[2] = 0.0;
          } else {
            const double tr = P[0] + P[4] + P[8];
            double Kb[9];
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
              for (int dd = 0; dd < 3; dd++)
                Kb[d * 3 + dd] = pl * P[d * 3 + dd] + pm * P[dd * 3 + d] + (d == dd ? pm * tr : 0.0);
            const double M = pr * mm;
            V[0] = Kb[0]; V[1] = Kb[1]; V[2] = Kb[2]; V[3] = Kb[3]; V[4] = Kb[4]; V[5] = M;
            V[6] = Kb[5]; V[7] = Kb[6]; V[8] = Kb[7]; V[9] = Kb[8]; V[10] = 0.0; V[11] = M; V[12] = 0.0;
          }
        }
        if (first) after_blocks();
        wave_sync();
        // 3. blocks to LDS, then the slot lanes add them in ascending element order
        if (k < nk) {
#pragma unroll
          for (int u = 0; u < Hex8Lds<ND>::kRec - 1; u++) L.u.v[lane][u] = V[u];
          if (slot >= s0 && slot < s0 + SPP) reinterpret_cast<uint8_t *>(&L.cb(slot - s0))[k] = (uint8_t)j;
        }
        wave_sync();
        if (s < c.deg) {
          const uint64_t c8 = L.cb(s - s0);
#pragma unroll
          for (int kk = 0; kk < 8; kk++) {
            const int bb = (int)((c8 >> (8 * kk)) & 0xFF);
            if (bb != 0xFF) {
              const double *src = L.u.v[kk * 8 + bb] + h * RH;
#pragma unroll
              for (int u = 0; u < RH; u++) acc[u] += src[u];
            }
          }
        }
      }
      // 4. the node's rows.  Common path: the slot lanes place their entries
      //    in an LDS image of the rows (slot t is columns ND t..ND t+ND-1),
      //    then the wavefront writes the rows with coalesced stores, a fixed
      //    number of them (lanes past the row end repeat its last entry with
      //    the same value) so that no store sits under a branch.
      if constexpr (!GENERAL) {
        constexpr int kCap = ND == 1 ? 64 : 96;  // row length of a node with one pass of slots
        static_assert(kCap <= 128, "two columns per lane");
        double *kr = L.u.rows;                   // [d * kCap + column]
        double *mr = L.u.rows + ND * kCap;       // mass per slot
        wave_sync();  // blocks and slot table read
        if (s < c.deg) {
          if constexpr (ND == 1) {
            kr[s] = acc[0];
            mr[s] = acc[1];
          } else {
#pragma unroll
            for (int u = 0; u < 5; u++) {
              const int en = h * 5 + u;  // entry (d, dd) of the 3x3 block, row-major
              if (en < 9) kr[(en / 3) * kCap + 3 * s + en % 3] = acc[u];
            }
            if (h == 0) mr[s] = acc[5];
          }
        }
        wave_sync();
        // two columns per lane (one 16-byte store each; rows hold at most
        // 2 x 64 entries here); lanes past the row end repeat its last pair
        const int c0 = min(2 * lane, c.len - 2);
#pragma unroll
        for (int d = 0; d < ND; d++) {
          const double k0 = kr[d * kCap + c0], k1 = kr[d * kCap + c0 + 1];
          const double m0 = ND == 1 ? mr[c0] : (c0 % 3 == d ? mr[c0 / 3] : 0.0);
          const double m1 = ND == 1 ? mr[c0 + 1] : ((c0 + 1) % 3 == d ? mr[(c0 + 1) / 3] : 0.0);
          const int64_t rb = (int64_t)c.rowbase + (int64_t)d * c.len + c0;
          *reinterpret_cast<double2 *>(Kv + rb) = make_double2(k0, k1);
          *reinterpret_cast<double2 *>(Mv + rb) = make_double2(m0, m1);
          if constexpr (WITH_C)
            *reinterpret_cast<double2 *>(Cv + rb) = make_double2(alpha * k0 + beta * m0, alpha * k1 + beta * m1);
        }
      } else if (s < c.deg) {
        // GENERAL: the slot lanes' entries straight into the rows; a slot that
        // no element touches (an interface partner of another field) holds zeros
        const int32_t wid = (s + 1 < c.deg ? offn : c.len) - off;
        if constexpr (ND == 1) {
          const int64_t rb = (int64_t)c.rowbase + off;
          if (wid == 1) {
            Kv[rb] = acc[0];
            Mv[rb] = acc[1];
            if constexpr (WITH_C) Cv[rb] = alpha * acc[0] + beta * acc[1];
          } else {
            for (int u = 0; u < wid; u++) {
              Kv[rb + u] = 0.0;
              Mv[rb + u] = 0.0;
              if constexpr (WITH_C) Cv[rb + u] = 0.0;
            }
          }
        } else {
          if (wid == 3) {
#pragma unroll
            for (int u = 0; u < 5; u++) {
              const int en = h * 5 + u;  // entry (d, dd) of the 3x3 block, row-major
              if (en < 9) {
                const int d = en / 3, dd = en - 3 * (en / 3);
                const int64_t rb = (int64_t)c.rowbase + (int64_t)d * c.len + off + dd;
                const double kv = acc[u], mv = d == dd ? acc[5] : 0.0;
                Kv[rb] = kv;
                Mv[rb] = mv;
                if constexpr (WITH_C) Cv[rb] = alpha * kv + beta * mv;
              }
            }
          } else if (h == 0) {
            for (int d = 0; d < 3; d++)
              for (int u = 0; u < wid; u++) {
                const int64_t rb = (int64_t)c.rowbase + (int64_t)d * c.len + off + u;
                Kv[rb] = 0.0;
                Mv[rb] = 0.0;
                if constexpr (WITH_C) Cv[rb] = 0.0;
              }
          }
        }
      }
    }
  };
  auto none = [] {};
  if constexpr (GENERAL) {
    for (; t < end; t += wx) {
      Hex8Fetch c{};
      fetch1(c, t);
      fetch2(c);
      fetch3(c);
      fetch4(c);
      process(c, none, none, none);
    }
  } else {
    // two nodes ahead: X = node t (all inputs in), Y = t + wx (levels 1-2 in,
    // 3-4 issued during t), Z = t + 2 wx (levels 1-2 issued during t); the
    // three records rotate, so nothing is copied while its loads are in flight
    // (past the end of the range the last node is fetched again, unused:
    // the loads stay unconditional, see fetch2)
    Hex8Fetch A{}, B{}, C{};
    fetch1(A, t);
    fetch2(A);
    fetch3(A);
    fetch4(A);
    fetch1(B, min(t + wx, end - 1));
    fetch2(B);
    auto step = [&](const Hex8Fetch &X, Hex8Fetch &Y, Hex8Fetch &Z) {
      process(
          X,
          [&] {
            fetch3(Y);
            fetch1(Z, min(t + 2 * wx, end - 1));
          },
          [&] { fetch4(Y); }, [&] { fetch2(Z); });
      t += wx;
      return t < end;
    };
    while (step(A, B, C) && step(B, C, A) && step(C, A, B)) {
    }
  }
  if (bad_elem != INT32_MAX) atomicMin(bad, bad_elem);
}

// persistent wavefronts for n descriptors: one per CU fewer than can be
// resident (a wave that does not find a free slot would run its whole share
// after the others), a multiple of the 8 XCDs
template <int ND, bool GENERAL, bool WITH_C>
static void launch_hex8_part(Model &m, KindBlock &kb, int64_t first, int64_t n, int *bad, hipStream_t s) {
  if (n == 0) return;
  int dev = 0, cus = 0;
  ELP_HIP(hipGetDevice(&dev));
  ELP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  static const int per_cu = [] {  // resident waves of this instantiation per CU (the same on every gfx950)
    int v = 0;
    ELP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, hex8_node_kernel<ND, GENERAL, WITH_C>, 64, 0));
    return v;
  }();
  const int64_t resident = (int64_t)std::max(cus, 8) * std::max(per_cu - 1, 1);
  const int grid = (int)(8 * std::max<int64_t>(1, std::min((n + 7) / 8, resident / 8)));
  hipLaunchKernelGGL((hex8_node_kernel<ND, GENERAL, WITH_C>), dim3(grid), dim3(64), 0, s, (int32_t)n,
                     kb.d_h8_desc.p + first * 8, kb.d_adj_elem.p, kb.d_adj_local.p, kb.d_adj_slot.p, kb.d_conn.p,
                     kb.d_params.p, m.d_xyz.p, m.d_nbr_off.p, m.d_K.p, m.d_M.p,
                     m.has_rayleigh ? m.d_C.p : nullptr, m.ray_a, m.ray_b, bad);
  ELP_HIP(hipGetLastError());
}

template <int ND, bool WITH_C>
static void launch_hex8_kind(Model &m, KindBlock &kb, int *bad, hipStream_t s) {
  const int64_t nn = (int64_t)kb.nodes.size(), ns = kb.h8_simple;
  launch_hex8_part<ND, false, WITH_C>(m, kb, 0, ns, bad, s);
  launch_hex8_part<ND, true, WITH_C>(m, kb, ns, nn - ns, bad, s);
}

static void launch_hex8_node(Model &m, KindBlock &kb, int *bad, hipStream_t s) {
  if (kb.kind == ELP_FLUID8) {
    if (m.has_rayleigh) launch_hex8_kind<1, true>(m, kb, bad, s);
    else launch_hex8_kind<1, false>(m, kb, bad, s);
  } else {
    if (m.has_rayleigh) launch_hex8_kind<3, true>(m, kb, bad, s);
    else launch_hex8_kind<3, false>(m, kb, bad, s);
  }
}

static bool fused_kind(const KindBlock &kb) { return kb.npe == 8; }

static void launch_volume(Model &m, KindBlock &kb, int *bad, hipStream_t s) {
  if (fused_kind(kb)) {
    launch_hex8_node(m, kb, bad, s);
  } else {
    launch_element(kb, m.d_xyz.p, 0, kb.d_Ke.p, kb.d_Me.p, bad, s, true);
    launch_gather(m, kb, s);
  }
}

void assemble_model(Model &m, hipStream_t s) {
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "elp_assemble before elp_model_finalize");
  Ctx *c = m.ctx;
  if (!m.d_K.p) {
    m.d_K.alloc(c, m.nnz);
    m.d_M.alloc(c, m.nnz);
    if (m.has_rayleigh) m.d_C.alloc(c, m.nnz);
    for (auto &kb : m.kinds) {
      if (fused_kind(kb)) continue;  // no element matrices in HBM
      // node-ordered records of K rows and M rows (element_kernel's ROWS layout)
      const int64_t ndf = (int64_t)kb.nd * kb.npe;
      kb.d_Ke.alloc(c, kb.n_elem * (ndf * ndf + (int64_t)kb.npe * kb.npe));
    }
    const int64_t fb = m.n_faces * m.nfn * m.nfn + m.n_if * 48;
    m.d_fbuf.alloc(c, fb);
    m.d_x_val.alloc(c, m.x_col.size());
  }
  DArr<int> bad;
  bad.alloc(nullptr, 2);
  std::vector<int> hb = {INT32_MAX, INT32_MAX};
  ELP_HIP(hipMemcpyAsync(bad.p, hb.data(), 2 * sizeof(int), hipMemcpyHostToDevice, s));
  for (auto &kb : m.kinds) launch_volume(m, kb, bad.p, s);
  if (m.n_faces) {
    if (m.nfn == 4)
      hipLaunchKernelGGL((face_kernel<4, 4>), dim3(blocks_for(m.n_faces, 128)), dim3(128), 0, s, m.n_faces,
                         m.d_face_nodes.p, m.d_xyz.p, m.d_fbuf.p, bad.p + 1);
    else
      hipLaunchKernelGGL((face_kernel<8, 9>), dim3(blocks_for(m.n_faces, 128)), dim3(128), 0, s, m.n_faces,
                         m.d_face_nodes.p, m.d_xyz.p, m.d_fbuf.p, bad.p + 1);
    ELP_HIP(hipGetLastError());
  }
  if (m.n_if) {
    hipLaunchKernelGGL(fsi_kernel, dim3(blocks_for(m.n_if, 128)), dim3(128), 0, s, m.n_if, m.d_if_fnodes.p,
                       m.d_xyz.p, m.d_if_factor.p, m.d_fbuf.p + m.n_faces * m.nfn * m.nfn);
    ELP_HIP(hipGetLastError());
  }
  const int64_t nx = (int64_t)m.x_col.size();
  if (nx) {
    hipLaunchKernelGGL(segmented_sum_kernel, dim3(blocks_for(nx, 256)), dim3(256), 0, s, nx, m.d_xc_ptr.p,
                       m.d_xc_idx.p, m.d_fbuf.p, m.d_x_val.p);
    ELP_HIP(hipGetLastError());
  }
  ELP_HIP(hipMemcpyAsync(hb.data(), bad.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  ELP_HIP(hipStreamSynchronize(s));
  ELP_REQUIRE(hb[0] == INT32_MAX, ELP_EDETJ, "element with detJ <= 0 (kind-local index " + std::to_string(hb[0]) + ")");
  ELP_REQUIRE(hb[1] == INT32_MAX, ELP_EDETJ, "impedance face with detJ <= 0 (face " + std::to_string(hb[1]) + ")");
  m.assembled = true;
}

float time_volume_assembly(Model &m, int repeats, hipStream_t s) {
  ELP_REQUIRE(m.assembled, ELP_ESTATE, "assemble first");
  DArr<int> bad;
  bad.alloc(nullptr, 1);
  HipEvent a, b;
  ELP_HIP(hipEventRecord(a.e, s));
  for (int r = 0; r < repeats; r++) {
    for (auto &kb : m.kinds) launch_volume(m, kb, bad.p, s);
  }
  ELP_HIP(hipEventRecord(b.e, s));
  ELP_HIP(hipEventSynchronize(b.e));
  float ms = 0;
  ELP_HIP(hipEventElapsedTime(&ms, a.e, b.e));
  return ms / repeats;
}

void element_matrices_batch(Ctx *ctx, int kind, int64_t n, const double *xyz, double p0, double p1,
                            double p2, double *Ke, double *Me) {
  ELP_REQUIRE(kind >= 1 && kind <= 4, ELP_EINVAL, "unknown element kind");
  ELP_REQUIRE(n >= 0, ELP_EINVAL, "negative element count");
  if (n == 0) return;
  KindBlock kb;
  kb.kind = kind;
  kb.npe = npe_of(kind);
  kb.nd = nd_of(kind);
  kb.n_elem = n;
  std::vector<double> par(n * 4);
  for (int64_t e = 0; e < n; e++) {
    if (is_fluid(kind)) {
      par[e * 4] = 1.0;           // no 1/rho here: the element operator itself
      par[e * 4 + 1] = 1.0 / (p1 * p1);
    } else {
      double f = p0 / (1.0 + p1) / (1.0 - 2.0 * p1);
      par[e * 4] = f * p1;
      par[e * 4 + 1] = f * (1.0 - 2.0 * p1) / 2.0;
      par[e * 4 + 2] = p2;
    }
  }
  kb.d_params.upload(ctx, par);
  DArr<double> dx;
  dx.alloc(ctx, n * kb.npe * 3);
  ELP_HIP(hipMemcpy(dx.p, xyz, dx.bytes(), hipMemcpyHostToDevice));
  const int64_t ndf = (int64_t)kb.nd * kb.npe;
  DArr<double> dK, dM;
  dK.alloc(ctx, n * ndf * ndf);
  dM.alloc(ctx, n * kb.npe * kb.npe);
  DArr<int> bad;
  bad.alloc(nullptr, 1);
  int hb = INT32_MAX;
  ELP_HIP(hipMemcpy(bad.p, &hb, sizeof(int), hipMemcpyHostToDevice));
  launch_element(kb, dx.p, 1, dK.p, dM.p, bad.p, ctx->stream);
  ELP_HIP(hipStreamSynchronize(ctx->stream));
  ELP_HIP(hipMemcpy(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost));
  ELP_REQUIRE(hb == INT32_MAX, ELP_EDETJ, "element with detJ <= 0 (index " + std::to_string(hb) + ")");
  ELP_HIP(hipMemcpy(Ke, dK.p, dK.bytes(), hipMemcpyDeviceToHost));
  if (Me) ELP_HIP(hipMemcpy(Me, dM.p, dM.bytes(), hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------
// elPaSo's system A(w) in the full pattern (for PETSc hand-off and parity)
// ---------------------------------------------------------------------------
struct Coefs {
  double kf_re, kf_im, mf_re, mf_im;  // fluid rows
  double ks_re, ks_im, ms_re, ms_im;  // structure rows
  double g_re, g_im;                  // C_Gamma
};

__global__ void dyn_rows_kernel(int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
                                const uint8_t *__restrict__ field, const double *__restrict__ K,
                                const double *__restrict__ M, Coefs c, double2 *__restrict__ A) {
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (r >= n) return;
  const uint8_t f = field[r];
  const double kr = f == 1 ? c.kf_re : c.ks_re, ki = f == 1 ? c.kf_im : c.ks_im;
  const double mr = f == 1 ? c.mf_re : c.ms_re, mi = f == 1 ? c.mf_im : c.ms_im;
  for (int32_t p = rowptr[r] + lane; p < rowptr[r + 1]; p += 64) {
    double2 v;
    if (f == 2) {
      v.x = (colind[p] == r) ? 1.0 : 0.0;
      v.y = 0.0;
    } else {
      v.x = kr * K[p] + mr * M[p];
      v.y = ki * K[p] + mi * M[p];
    }
    A[p] = v;
  }
}

__global__ void dyn_extra_kernel(int64_t n, const int32_t *__restrict__ x_rowptr, const int32_t *__restrict__ fullpos,
                                 const uint8_t *__restrict__ xtype, const double *__restrict__ xval,
                                 const uint8_t *__restrict__ field, Coefs c, double2 *__restrict__ A) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || field[r] == 2) return;
  for (int32_t k = x_rowptr[r]; k < x_rowptr[r + 1]; k++) {
    double2 v = A[fullpos[k]];
    if (xtype[k] == 0) {
      v.x += c.g_re * xval[k];
      v.y += c.g_im * xval[k];
    } else {
      v.x += xval[k];
    }
    A[fullpos[k]] = v;
  }
}

void dynamic_matrix(Model &m, double w, std::complex<double> Z, double *a, double *f) {
  ELP_REQUIRE(m.assembled, ELP_ESTATE, "assemble first");
  ELP_REQUIRE(w > 0, ELP_EINVAL, "omega must be positive");
  const std::complex<double> I(0, 1);
  std::complex<double> kf = 1.0 / (w * w), mf = -1.0;
  std::complex<double> ks = 1.0 + I * m.eta * (m.visco_type == 2 ? w : 1.0);
  std::complex<double> ms = -w * w;
  if (m.has_rayleigh) {
    ks += I * w * m.ray_a;
    ms += I * w * m.ray_b;
  }
  std::complex<double> g = (m.n_faces && std::abs(Z) > 0) ? I / (w * Z) : 0.0;
  Coefs c{kf.real(), kf.imag(), mf.real(), mf.imag(), ks.real(), ks.imag(), ms.real(), ms.imag(), g.real(), g.imag()};
  hipStream_t s = m.ctx->stream;
  DArr<double2> dA;
  dA.alloc(m.ctx, m.nnz);
  hipLaunchKernelGGL(dyn_rows_kernel, dim3(blocks_for(m.n, 4)), dim3(256), 0, s, m.n, m.d_rowptr.p, m.d_colind.p,
                     m.d_row_field.p, m.d_K.p, m.d_M.p, c, dA.p);
  ELP_HIP(hipGetLastError());
  if (!m.x_col.empty()) {
    hipLaunchKernelGGL(dyn_extra_kernel, dim3(blocks_for(m.n, 256)), dim3(256), 0, s, m.n, m.d_x_rowptr.p,
                       m.d_x_fullpos.p, m.d_x_type.p, m.d_x_val.p, m.d_row_field.p, c, dA.p);
    ELP_HIP(hipGetLastError());
  }
  ELP_HIP(hipStreamSynchronize(s));
  if (a) ELP_HIP(hipMemcpy(a, dA.p, m.nnz * sizeof(double2), hipMemcpyDeviceToHost));
  if (f) {
    for (int64_t r = 0; r < m.n; r++) {
      f[2 * r] = m.F[r].real();
      f[2 * r + 1] = m.F[r].imag();
    }
    for (size_t k = 0; k < m.dir_rows.size(); k++) {
      f[2 * m.dir_rows[k]] = m.dir_vals[k].real();
      f[2 * m.dir_rows[k] + 1] = m.dir_vals[k].imag();
    }
  }
}

}  // namespace elp
