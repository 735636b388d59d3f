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

// Element matrices of fluid elements: K_ab = sum_g wdJ grad_a.grad_b / rho,
// M_ab = sum_g (1/c^2) N_a N_b wdJ / rho  (elementfluid3d.cpp:54-167, with the
// frequency-independent 1/rho of elPaSo's 1/(rho w^2) scaling, elementfluid.cpp:74-76).
// Structural elements: K = sum_g B^T D B wdJ for isotropic D in closed form
// lambda p_d q_e + mu p_e q_d + mu delta_de p.q, M_ab = rho sum_g N_a N_b wdJ
// (elementstructurebrick.cpp:470-554).  params: fluid (rho, 1/c^2), brick
// (lambda, mu, rho).  bad receives the smallest element index with detJ <= 0.
template <int NPE, int ND, int NG>
__global__ __launch_bounds__(64 * kWaves) void element_kernel(
    const double *__restrict__ xyz, const int32_t *__restrict__ conn, const double *__restrict__ params,
    int64_t n_elem, double *__restrict__ Ke, double *__restrict__ Me, int *bad, int xyz_direct) {
  __shared__ ElemLds<NPE, NG> lds[kWaves];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t e = (int64_t)blockIdx.x * kWaves + w;
  if (e >= n_elem) return;  // whole wavefront leaves together
  ElemLds<NPE, NG> &L = lds[w];
  for (int t = lane; t < NPE * 3; t += 64) {
    int k = t / 3, d = t % 3;
    L.x[t] = xyz_direct ? xyz[e * NPE * 3 + t] : xyz[(int64_t)conn[e * NPE + k] * 3 + d];
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
    for (int t = lane; t < NPE * NPE; t += 64) {
      const int a = t / NPE, b = t % NPE;
      double k = 0.0, m = 0.0;
#pragma unroll 3
      for (int g = 0; g < NG; g++) {
        const double wd = L.wdj[g];
        double r = L.grad[g][0][a] * L.grad[g][0][b] + L.grad[g][1][a] * L.grad[g][1][b] +
                   L.grad[g][2][a] * L.grad[g][2][b];
        k += r * wd;
        const double *N = tab_n<NPE>(g, 0);
        m += p1 * N[a] * N[b] * wd;
      }
      Kout[t] = k / p0;
      Mout[t] = m / p0;
    }
  } else {
    const double lam = p0, mu = p1, rho = p2;
    for (int t = lane; t < NDF * NDF; t += 64) {
      const int row = t / NDF, col = t % NDF;
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
      Kout[t] = k;
    }
    for (int t = lane; t < NPE * NPE; t += 64) {
      const int a = t / NPE, b = t % NPE;
      double m = 0.0;
#pragma unroll 3
      for (int g = 0; g < NG; g++) {
        const double *N = tab_n<NPE>(g, 0);
        m += rho * N[a] * N[b] * L.wdj[g];
      }
      Mout[t] = m;
    }
  }
}

// Deterministic gather of element rows into the node's CSR rows (K, M and
// optionally C = alpha K + beta M), one wavefront per node.
constexpr int kMaxCols = 256;
template <int NPE, int ND>
struct GatherLds {
  double k[ND][kMaxCols];
  double m[kMaxCols];
  int32_t nb[kMaxCols];
  int32_t off[kMaxCols + 1];
};

__device__ inline int find_slot(const int32_t *nb, int deg, int32_t j) {
  int lo = 0, hi = deg - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (nb[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <int NPE, int ND>
__global__ __launch_bounds__(64 * kWaves) void gather_kernel(
    int32_t n_nodes, const int32_t *__restrict__ nodes, const int32_t *__restrict__ adj_ptr,
    const int32_t *__restrict__ adj_elem, const uint8_t *__restrict__ adj_local,
    const int32_t *__restrict__ conn, const double *__restrict__ Ke, const double *__restrict__ Me,
    const int32_t *__restrict__ nbr_ptr, const int32_t *__restrict__ nbr,
    const int32_t *__restrict__ nbr_off, const int32_t *__restrict__ node_row,
    const int32_t *__restrict__ rowptr, double *__restrict__ Kv, double *__restrict__ Mv,
    double *__restrict__ Cv, double alpha, double beta) {
  __shared__ GatherLds<NPE, ND> lds[kWaves];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int32_t s_node = blockIdx.x * kWaves + w;
  if (s_node >= n_nodes) return;
  GatherLds<NPE, ND> &L = lds[w];
  const int32_t i = nodes[s_node];
  const int32_t nb0 = nbr_ptr[i], deg = nbr_ptr[i + 1] - nb0;
  for (int s = lane; s < deg; s += 64) {
    L.nb[s] = nbr[nb0 + s];
    L.off[s] = nbr_off[nb0 + s];
  }
  const int32_t r0 = node_row[i];
  const int32_t len = rowptr[r0 + 1] - rowptr[r0];
  if (lane == 0) L.off[deg] = len;
  for (int q = lane; q < len; q += 64) {
#pragma unroll
    for (int d = 0; d < ND; d++) L.k[d][q] = 0.0;
  }
  for (int s = lane; s < deg; s += 64) L.m[s] = 0.0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  constexpr int NDF = NPE * ND;
  for (int32_t p = adj_ptr[s_node]; p < adj_ptr[s_node + 1]; p++) {
    const int64_t e = adj_elem[p];
    const int a = adj_local[p];
    const double *kr = Ke + e * (int64_t)NDF * NDF + (int64_t)a * ND * NDF;
    for (int t = lane; t < ND * NDF; t += 64) {
      const int d = t / NDF, c = t % NDF, b = c / ND, dd = c % ND;
      const int s = find_slot(L.nb, deg, conn[e * NPE + b]);
      L.k[d][L.off[s] + dd] += kr[d * NDF + c];
    }
    if (lane < NPE) {
      const int s = find_slot(L.nb, deg, conn[e * NPE + lane]);
      L.m[s] += Me[e * NPE * NPE + a * NPE + lane];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  for (int d = 0; d < ND; d++) {
    const int64_t base = rowptr[r0 + d];
    for (int q = lane; q < len; q += 64) {
      // slot of column q and the DOF within the neighbour
      const int s = [&] {
        int lo = 0, hi = deg - 1;
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (L.off[mid] <= q) lo = mid;
          else hi = mid - 1;
        }
        return lo;
      }();
      const int dd = q - L.off[s];
      const double kv = L.k[d][q];
      const double mv = (dd == d) ? L.m[s] : 0.0;
      Kv[base + q] = kv;
      Mv[base + q] = mv;
      if (Cv) Cv[base + q] = alpha * kv + beta * mv;
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

static void launch_element(const KindBlock &kb, const double *xyz, int xyz_direct, double *Ke,
                           double *Me, int *bad, hipStream_t s) {
  const int grid = blocks_for(kb.n_elem, kWaves);
  if (grid == 0) return;
  const int32_t *conn = kb.d_conn.p;
  const double *par = kb.d_params.p;
  switch (kb.kind) {
    case ELP_FLUID8:
      hipLaunchKernelGGL((element_kernel<8, 1, 8>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                         kb.n_elem, Ke, Me, bad, xyz_direct);
      break;
    case ELP_BRICK8:
      hipLaunchKernelGGL((element_kernel<8, 3, 8>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                         kb.n_elem, Ke, Me, bad, xyz_direct);
      break;
    case ELP_FLUID20:
      hipLaunchKernelGGL((element_kernel<20, 1, 27>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                         kb.n_elem, Ke, Me, bad, xyz_direct);
      break;
    default:
      hipLaunchKernelGGL((element_kernel<20, 3, 27>), dim3(grid), dim3(64 * kWaves), 0, s, xyz, conn, par,
                         kb.n_elem, Ke, Me, bad, xyz_direct);
      break;
  }
  ELP_HIP(hipGetLastError());
}

static void launch_gather(Model &m, KindBlock &kb, hipStream_t s) {
  const int32_t nn = (int32_t)kb.nodes.size();
  const int grid = blocks_for(nn, kWaves);
  if (grid == 0) return;
  double *Cv = m.has_rayleigh ? m.d_C.p : nullptr;
#define ELP_GATHER_ARGS                                                                          \
  nn, kb.d_nodes.p, kb.d_adj_ptr.p, kb.d_adj_elem.p, kb.d_adj_local.p, kb.d_conn.p, kb.d_Ke.p,   \
      kb.d_Me.p, m.d_nbr_ptr.p, m.d_nbr.p, m.d_nbr_off.p, m.d_node_row.p, m.d_rowptr.p, m.d_K.p, \
      m.d_M.p, Cv, m.ray_a, m.ray_b
  switch (kb.kind) {
    case ELP_FLUID8:
      hipLaunchKernelGGL((gather_kernel<8, 1>), dim3(grid), dim3(64 * kWaves), 0, s, ELP_GATHER_ARGS);
      break;
    case ELP_BRICK8:
      hipLaunchKernelGGL((gather_kernel<8, 3>), dim3(grid), dim3(64 * kWaves), 0, s, ELP_GATHER_ARGS);
      break;
    case ELP_FLUID20:
      hipLaunchKernelGGL((gather_kernel<20, 1>), dim3(grid), dim3(64 * kWaves), 0, s, ELP_GATHER_ARGS);
      break;
    default:
      hipLaunchKernelGGL((gather_kernel<20, 3>), dim3(grid), dim3(64 * kWaves), 0, s, ELP_GATHER_ARGS);
      break;
  }
#undef ELP_GATHER_ARGS
  ELP_HIP(hipGetLastError());
}

// Hex8 shape values at a 2x2x2 Gauss point computed in registers, the same
// arithmetic as the host table (lin1 factors 1 +- 0.577350269189626 chosen
// by the signs of node and point, products ((a*b)*c)*0.125): bit-identical to
// c_tab.h8 without table loads.  Node k's natural coordinate signs are bits
// of 0x66 (x), 0xCC (y), 0xF0 (z); point g has x index g % 2, y (g / 2) % 2,
// z g / 4, index 0 = +0.577350269189626.
struct Hex8Shape {
  double n, dx, dy, dz;
};
__device__ inline Hex8Shape hex8_shape(int k, int g) {
  constexpr double gp = 0.577350269189626, hi = 1. + gp, lo = 1. - gp;
  const bool sx = (0x66 >> k) & 1, sy = (0xCC >> k) & 1, sz = (0xF0 >> k) & 1;
  const double a = (sx == !(g & 1)) ? hi : lo;
  const double b = (sy == !((g >> 1) & 1)) ? hi : lo;
  const double c = (sz == !((g >> 2) & 1)) ? hi : lo;
  Hex8Shape r;
  r.n = a * b * c * 0.125;
  r.dx = (sx ? 1.0 : -1.0) * (b * c * 0.125);
  r.dy = (sy ? 1.0 : -1.0) * (a * c * 0.125);
  r.dz = (sz ? 1.0 : -1.0) * (a * b * 0.125);
  return r;
}

// Fused assembly of trilinear (Hex8) elements directly into the CSR rows of
// one node per wavefront, without element matrices in HBM:
//  1. lane (k, g): geometry of adjacent element k at quadrature point g
//     (inverse Jacobian and w detJ) into LDS;
//  2. lane (k, b): the contribution of element k to the (node, neighbour b)
//     block, contracted over the 8 points from LDS;
//  3. lane per neighbour slot: sum of its contributions in ascending element
//     order (elPaSo's MatSetValues order), finished into K, M (and C) rows in
//     LDS;
//  4. the node's rows are written with coalesced stores.
constexpr int kMaxHex8Cols = 128;  // row length / neighbours of a Hex8 node (checked at finalize)
template <int ND>
struct Hex8Lds {
  // [(k*8+g)*10 + k + t]: invJ (t < 9), w detJ (t = 9); the pad of one
  // double per element puts the 8 elements' rows read together in phase 2
  // in distinct LDS banks
  double geo[64 * 10 + 8];
  union {
    double xe[64][3];  // phase 0-1: coordinates of node q of element k, [k*8+q]
    double kr[ND][kMaxHex8Cols];
  } u;
  double mr[kMaxHex8Cols];
  int32_t nb[kMaxHex8Cols];
  int32_t off[kMaxHex8Cols + 1];
  int16_t colslot[kMaxHex8Cols];
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int ND>
__global__ __launch_bounds__(64) void hex8_node_kernel(
    int32_t n_nodes, const int32_t *__restrict__ nodes, const int32_t *__restrict__ adj_ptr,
    const int32_t *__restrict__ adj_elem, const uint8_t *__restrict__ adj_local, const int32_t *__restrict__ conn,
    const double *__restrict__ params, const double *__restrict__ xyz, const int32_t *__restrict__ nbr_ptr,
    const int32_t *__restrict__ nbr, const int32_t *__restrict__ nbr_off, const int32_t *__restrict__ node_row,
    const int32_t *__restrict__ rowptr, double *__restrict__ Kv, double *__restrict__ Mv, double *__restrict__ Cv,
    double alpha, double beta, int *bad) {
  __shared__ Hex8Lds<ND> L;
  const int lane = threadIdx.x;
  const int k = lane / 8, g = lane % 8, b = lane % 8;
  {
    // consecutive nodes (sharing elements and coordinates) on one XCD: the
    // hardware deals workgroups round-robin to the 8 XCDs
    const int32_t per = gridDim.x / 8;
    const int32_t s_node = (int32_t)(blockIdx.x % 8) * per + (int32_t)(blockIdx.x / 8);
    if (s_node >= n_nodes) return;
    const int32_t i = nodes[s_node];
    const int32_t a0 = adj_ptr[s_node], nadj = adj_ptr[s_node + 1] - a0;
    const int32_t nb0 = nbr_ptr[i], deg = nbr_ptr[i + 1] - nb0;
    for (int t = lane; t < deg; t += 64) {
      L.nb[t] = nbr[nb0 + t];
      L.off[t] = nbr_off[nb0 + t];
    }
    const int32_t r0 = node_row[i];
    const int32_t len = rowptr[r0 + 1] - rowptr[r0];
    if (lane == 0) L.off[deg] = len;
    // 0. coordinates: lane (k, q) loads node q of element k
    int32_t e = -1, a = 0;
    if (k < nadj) {
      e = adj_elem[a0 + k];
      a = adj_local[a0 + k];
      const int32_t nd_ = conn[e * 8 + g];
      L.u.xe[lane][0] = xyz[nd_ * 3];
      L.u.xe[lane][1] = xyz[nd_ * 3 + 1];
      L.u.xe[lane][2] = xyz[nd_ * 3 + 2];
    }
    wave_sync();
    // 1. geometry of (element k, point g): inverse Jacobian, w detJ, and
    //    w detJ grad N_a of the node itself
    if (k < nadj) {
      double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const double *x = L.u.xe[k * 8 + q];
        const double x0 = x[0], x1 = x[1], x2 = x[2];
        const Hex8Shape h = hex8_shape(q, g);
        J[0] += h.dx * x0; J[1] += h.dx * x1; J[2] += h.dx * x2;
        J[3] += h.dy * x0; J[4] += h.dy * x1; J[5] += h.dy * x2;
        J[6] += h.dz * x0; J[7] += h.dz * x1; J[8] += h.dz * x2;
      }
      const double det = J[0] * J[4] * J[8] + J[1] * J[5] * J[6] + J[2] * J[3] * J[7] -
                         J[2] * J[4] * J[6] - J[0] * J[5] * J[7] - J[1] * J[3] * J[8];
      if (!(det > 0.0)) atomicMin(bad, e);
      const double id = 1.0 / det;
      double G[9];
      G[0] = (J[4] * J[8] - J[5] * J[7]) * id;
      G[3] = -(J[3] * J[8] - J[5] * J[6]) * id;
      G[6] = (J[3] * J[7] - J[4] * J[6]) * id;
      G[1] = -(J[1] * J[8] - J[2] * J[7]) * id;
      G[4] = (J[0] * J[8] - J[2] * J[6]) * id;
      G[7] = -(J[0] * J[7] - J[1] * J[6]) * id;
      G[2] = (J[1] * J[5] - J[2] * J[4]) * id;
      G[5] = -(J[0] * J[5] - J[2] * J[3]) * id;
      G[8] = (J[0] * J[4] - J[1] * J[3]) * id;
      double *O = L.geo + lane * 10 + k;
#pragma unroll
      for (int t = 0; t < 9; t++) O[t] = G[t];
      O[9] = c_tab.w8[g] * det;
    }
    wave_sync();
    // zero the row accumulators (they reuse the coordinate staging) and map
    // every column of the rows to its neighbour slot
    for (int q = lane; q < len; q += 64) {
#pragma unroll
      for (int d = 0; d < ND; d++) L.u.kr[d][q] = 0.0;
    }
    for (int sl = lane; sl < deg; sl += 64) {
      L.mr[sl] = 0.0;
      for (int q = L.off[sl]; q < L.off[sl + 1]; q++) L.colslot[q] = (int16_t)sl;
    }
    // 2. unit (k, b): element k's block of the (node, neighbour b) pair, with
    //    the element's own material law
    double V[ND == 1 ? 2 : 10];
    int slot = -1;
    if (k < nadj) {
      slot = find_slot(L.nb, deg, conn[e * 8 + b]);
      double P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, mm = 0.0, kk = 0.0;
#pragma unroll
      for (int gg = 0; gg < 8; gg++) {
        const double *G = L.geo + (k * 8 + gg) * 10 + k;
        const Hex8Shape ha = hex8_shape(a, gg), hb = hex8_shape(b, gg);
        const double ax = ha.dx, ay = ha.dy, az = ha.dz;
        const double bx = hb.dx, by = hb.dy, bz = hb.dz;
        const double wd = G[9];
        const double p0 = G[0] * ax + G[1] * ay + G[2] * az;
        const double p1 = G[3] * ax + G[4] * ay + G[5] * az;
        const double p2 = G[6] * ax + G[7] * ay + G[8] * az;
        const double q0 = G[0] * bx + G[1] * by + G[2] * bz;
        const double q1 = G[3] * bx + G[4] * by + G[5] * bz;
        const double q2 = G[6] * bx + G[7] * by + G[8] * bz;
        mm += ha.n * hb.n * wd;
        const double w0 = wd * p0, w1 = wd * p1, w2 = wd * p2;
        if constexpr (ND == 1) {
          kk += w0 * q0 + w1 * q1 + w2 * q2;
        } else {
          P[0] += w0 * q0; P[1] += w0 * q1; P[2] += w0 * q2;
          P[3] += w1 * q0; P[4] += w1 * q1; P[5] += w1 * q2;
          P[6] += w2 * q0; P[7] += w2 * q1; P[8] += w2 * q2;
        }
      }
      const double p0m = params[e * 4], p1m = params[e * 4 + 1], p2m = params[e * 4 + 2];
      if constexpr (ND == 1) {
        V[0] = kk / p0m;
        V[1] = p1m * mm / p0m;
      } else {
        const double tr = P[0] + P[4] + P[8];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
          for (int dd = 0; dd < 3; dd++)
            V[d * 3 + dd] = p0m * P[d * 3 + dd] + p1m * P[dd * 3 + d] + (d == dd ? p1m * tr : 0.0);
        V[9] = p2m * mm;
      }
    }
    wave_sync();
    // 3. ordered accumulation: element k's 8 units (distinct slots) add in
    //    round k, so every entry sums in ascending element id
    for (int kk = 0; kk < nadj; kk++) {
      if (k == kk) {
        const int off = L.off[slot];
        if constexpr (ND == 1) {
          L.u.kr[0][off] += V[0];
          L.mr[slot] += V[1];
        } else {
#pragma unroll
          for (int d = 0; d < 3; d++)
#pragma unroll
            for (int dd = 0; dd < 3; dd++) L.u.kr[d][off + dd] += V[d * 3 + dd];
          L.mr[slot] += V[9];
        }
      }
      wave_sync();
    }
    // 4. coalesced row writes
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const int64_t base = rowptr[r0 + d];
      for (int q = lane; q < len; q += 64) {
        const int sl = L.colslot[q];
        const int dd = q - L.off[sl];
        const double kv = L.u.kr[d][q];
        const double mv = (dd == d) ? L.mr[sl] : 0.0;
        Kv[base + q] = kv;
        Mv[base + q] = mv;
        if (Cv) Cv[base + q] = alpha * kv + beta * mv;
      }
    }
  }
}

static void launch_hex8_node(Model &m, KindBlock &kb, int *bad, hipStream_t s) {
  const int32_t nn = (int32_t)kb.nodes.size();
  if (nn == 0) return;
  const int grid = (nn + 7) / 8 * 8;  // one wavefront per node, a multiple of the 8 XCDs
  double *Cv = m.has_rayleigh ? m.d_C.p : nullptr;
#define ELP_HEX8_ARGS                                                                                   \
  nn, kb.d_nodes.p, kb.d_adj_ptr.p, kb.d_adj_elem.p, kb.d_adj_local.p, kb.d_conn.p, kb.d_params.p,   \
      m.d_xyz.p, m.d_nbr_ptr.p, m.d_nbr.p, m.d_nbr_off.p, m.d_node_row.p, m.d_rowptr.p, m.d_K.p,     \
      m.d_M.p, Cv, m.ray_a, m.ray_b, bad
  if (kb.kind == ELP_FLUID8)
    hipLaunchKernelGGL(hex8_node_kernel<1>, dim3(grid), dim3(64), 0, s, ELP_HEX8_ARGS);
  else
    hipLaunchKernelGGL(hex8_node_kernel<3>, dim3(grid), dim3(64), 0, s, ELP_HEX8_ARGS);
#undef ELP_HEX8_ARGS
  ELP_HIP(hipGetLastError());
}

static bool fused_kind(const KindBlock &kb) { return kb.npe == 8; }

static void launch_volume(Model &m, KindBlock &kb, int *bad, hipStream_t s) {
  if (fused_kind(kb)) {
    launch_hex8_node(m, kb, bad, s);
  } else {
    launch_element(kb, m.d_xyz.p, 0, kb.d_Ke.p, kb.d_Me.p, bad, s);
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
      const int64_t ndf = (int64_t)kb.nd * kb.npe;
      kb.d_Ke.alloc(c, kb.n_elem * ndf * ndf);
      kb.d_Me.alloc(c, kb.n_elem * (int64_t)kb.npe * kb.npe);
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
  hipEvent_t a, b;
  ELP_HIP(hipEventCreate(&a));
  ELP_HIP(hipEventCreate(&b));
  ELP_HIP(hipEventRecord(a, s));
  for (int r = 0; r < repeats; r++) {
    for (auto &kb : m.kinds) launch_volume(m, kb, bad.p, s);
  }
  ELP_HIP(hipEventRecord(b, s));
  ELP_HIP(hipEventSynchronize(b));
  float ms = 0;
  ELP_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
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
