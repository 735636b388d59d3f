// Results at elements and nodes from a block of solutions, on the device:
// the pressure gradient of the fluid elements at their nodes
// (cElementFluid3d::computeStressesNodes / computeGradP / computeGradPPoint,
// element/fluid/elementfluid3d.cpp:380-540) and the stresses of the bricks at
// their centroid (cElementStructureBrick::computeStressesCells /
// computeStressesNodes, element/structure/linear/brick/elementstructurebrick.cpp:48-279),
// collected per node as cAnalysis::computeStressesNodes does
// (analysis/analysis.cpp:869-937): 9 values per node (sigma 0..5, grad p 6..8).
//
// Layout: solutions X[f][row] (complex), f over a chunk of frequencies.  Per
// chunk: one thread per (brick, frequency) for the centroid stresses, one per
// (node, frequency) summing them over the node's bricks, and one per (fluid
// node, element) applying the element's gradient at the node to all of the
// chunk's frequencies.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "elp_fem.h"
#include "elp_internal.h"

namespace elp {

namespace {

constexpr int kPostThreads = 128;

// J(i,j) = sum_k dN_k/dxi_i x_k[j]; det by Sarrus' rule, inverse by cofactors
// (element/elementfem.cpp:158-210)
__device__ inline double jac_inverse(const double J[3][3], double inv[3][3]) {
  const double det = J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
                     J[0][2] * J[1][1] * J[2][0] - J[0][0] * J[1][2] * J[2][1] - J[0][1] * J[1][0] * J[2][2];
  const double r = 1.0 / det;
  inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * r;
  inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
  inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * r;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * r;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
  return det;
}

// grad p at the fluid nodes: one thread per node of a fluid kind (kinds that
// fluid_item_gradp_kernel below does not take); for each
// of the node's elements in ascending id the element's gradient operator at
// the node (J^-1 dN, frequency independent) is formed once and applied to the
// chunk's nf solutions, each added to the node's slots 6..8 of out[f][r][9]
// (the node's own slots: no other thread writes them).  elPaSo sums the NPE
// terms of a component sorted by magnitude (std::sort with abs_order,
// misc/mytypes.h:226, elementfluid3d.cpp:521-526); here they are summed in
// the element's local node order with fused multiply-adds: the two differ by
// rounding only (a few ulps of the largest term).
template <int NPE>
__global__ __launch_bounds__(kPostThreads) void fluid_node_gradp_kernel(
    int64_t n_kn, const int32_t *__restrict__ knodes, const int32_t *__restrict__ adj_ptr,
    const int32_t *__restrict__ adj_elem, const uint8_t *__restrict__ adj_local, const int32_t *__restrict__ conn,
    const double *__restrict__ xyz, const int32_t *__restrict__ node_row, const double *__restrict__ T, int nf,
    int64_t n, int64_t n_nodes, const double2 *__restrict__ X, double2 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_kn) return;
  const int32_t r = knodes[i];
  for (int32_t j = adj_ptr[i]; j < adj_ptr[i + 1]; j++) {
    const int64_t e = adj_elem[j];
    const double *Ta = T + (int)adj_local[j] * 3 * NPE;
    double x[NPE][3];
    int32_t row[NPE];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
      const int32_t q = conn[e * NPE + k];
      x[k][0] = xyz[(int64_t)q * 3];
      x[k][1] = xyz[(int64_t)q * 3 + 1];
      x[k][2] = xyz[(int64_t)q * 3 + 2];
      row[k] = node_row[q];
    }
    double J[3][3], inv[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NPE; k++) s += Ta[a * NPE + k] * x[k][b];
        J[a][b] = s;
      }
    jac_inverse(J, inv);
    double jh[3][NPE];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int k = 0; k < NPE; k++)
        jh[a][k] = inv[a][0] * Ta[k] + inv[a][1] * Ta[NPE + k] + inv[a][2] * Ta[2 * NPE + k];
    for (int f = 0; f < nf; f++) {
      const double2 *Xf = X + (int64_t)f * n;
      double2 p[NPE];
#pragma unroll
      for (int k = 0; k < NPE; k++) p[k] = Xf[row[k]];
      double2 *o = out + ((int64_t)f * n_nodes + r) * 9 + 6;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        double2 g = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NPE; k++) {
          g.x = fma(jh[a][k], p[k].x, g.x);
          g.y = fma(jh[a][k], p[k].y, g.y);
        }
        double2 v = o[a];
        v.x += g.x;
        v.y += g.y;
        o[a] = v;
      }
    }
  }
}

// grad p at the fluid nodes, one thread per (node, element) entry of the
// kind's adjacency.  The kind's nodes are cut into runs of whole nodes with
// at most kGradItems entries (run r: entries blk[r] .. blk[r + 1] of
// `knodes`), one block of kGradItems threads each.  Each thread forms its
// element's inverse Jacobian at the node once (the table T of shape-function
// derivatives at the element's nodes staged in LDS) and keeps it and the
// byte offsets of the element's rows across the chunk's frequencies; per
// frequency it gathers the element's nodal pressures, takes the derivatives
// along the element's axes, r_c = sum_k dN_k/dxi_c p_k, in local node order
// and maps them to grad p = J^-1 r (one FMA chain per component;
// J(c, d) = dx_d / dxi_c).  Thread t < (nodes of the run) then adds the
// gradients of node t's elements to its slots 6..8 of out in ascending
// element id (loaded before the gathers, so that their latency overlaps), as
// the per-node kernel above does; the two differ only in the association of
// the same products.  Needs n * 16 < 2^31 (32-bit buffer offsets into a
// frequency's solution).
constexpr int kGradItems = 256;
template <int NPE>
__global__ __launch_bounds__(kGradItems) void fluid_item_gradp_kernel(
    const int32_t *__restrict__ blk, const int32_t *__restrict__ knodes, const int32_t *__restrict__ adj_ptr,
    const int32_t *__restrict__ adj_elem, const uint8_t *__restrict__ adj_local, const int32_t *__restrict__ conn,
    const double *__restrict__ xyz, const int32_t *__restrict__ node_row, const double *__restrict__ T, int nf,
    int64_t n, int64_t n_nodes, const double2 *__restrict__ X, double2 *__restrict__ out) {
  constexpr int kTa = 3 * NPE + 1;
  __shared__ double sT[NPE * kTa];
  __shared__ double2 sG[kGradItems][3];
  __shared__ int32_t sSeg[kGradItems + 1];
  const int t = threadIdx.x;
  const int32_t nb0 = blk[blockIdx.x], nb1 = blk[blockIdx.x + 1];
  const int32_t ip0 = adj_ptr[nb0], ni = adj_ptr[nb1] - ip0;
  for (int u = t; u < NPE * 3 * NPE; u += kGradItems) sT[(u / (3 * NPE)) * kTa + u % (3 * NPE)] = T[u];
  for (int u = t; u <= nb1 - nb0; u += kGradItems) sSeg[u] = adj_ptr[nb0 + u] - ip0;
  __syncthreads();
  const bool has = t < ni;
  const int32_t j = ip0 + min(t, max(ni - 1, 0));
  const int64_t e = adj_elem[j];
  const double *Ta = sT + (int)adj_local[j] * kTa;
  uint32_t ro[NPE];
  double inv[3][3];
  {
    double J[3][3] = {};
#pragma unroll
    for (int k = 0; k < NPE; k++) {
      const int32_t q = conn[e * NPE + k];
      const double x0 = xyz[(int64_t)q * 3], x1 = xyz[(int64_t)q * 3 + 1], x2 = xyz[(int64_t)q * 3 + 2];
      ro[k] = (uint32_t)node_row[q] * 16u;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double d = Ta[c * NPE + k];
        J[c][0] += d * x0;
        J[c][1] += d * x1;
        J[c][2] += d * x2;
      }
    }
    jac_inverse(J, inv);
  }
  const int32_t nn = nb1 - nb0;
  typedef double d2v __attribute__((ext_vector_type(2)));
  for (int f = 0; f < nf; f++) {
    const __amdgpu_buffer_rsrc_t xr =
        __builtin_amdgcn_make_buffer_rsrc((void *)(X + (int64_t)f * n), (short)0, (int)(n * 16), 0x00020000);
    double2 *o = out + ((int64_t)f * n_nodes + knodes[nb0 + min(t, nn - 1)]) * 9 + 6;
    double2 v[3] = {};
    if (t < nn) {
#pragma unroll
      for (int a = 0; a < 3; a++) v[a] = o[a];
    }
    if (has) {
      double2 r[3] = {};
#pragma unroll
      for (int k = 0; k < NPE; k++) {
        const d2v p = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(xr, ro[k], 0, 0));
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double d = Ta[c * NPE + k];
          r[c].x = fma(d, p.x, r[c].x);
          r[c].y = fma(d, p.y, r[c].y);
        }
      }
#pragma unroll
      for (int a = 0; a < 3; a++) {
        double2 g;
        g.x = fma(inv[a][2], r[2].x, fma(inv[a][1], r[1].x, inv[a][0] * r[0].x));
        g.y = fma(inv[a][2], r[2].y, fma(inv[a][1], r[1].y, inv[a][0] * r[0].y));
        sG[t][a] = g;
      }
    }
    __syncthreads();
    if (t < nn) {
      for (int32_t u = sSeg[t]; u < sSeg[t + 1]; u++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
          v[a].x += sG[u][a].x;
          v[a].y += sG[u][a].y;
        }
#pragma unroll
      for (int a = 0; a < 3; a++) o[a] = v[a];
    }
    __syncthreads();
  }
}

// stresses of one brick at its centroid from its 8 corner nodes:
// out[f][e][6] = scale_f * D * B u (strain order xx, yy, zz, xy, yz, xz as
// the B matrix of elementstructurebrick.cpp:97-107)
__global__ __launch_bounds__(kPostThreads) void brick_centroid_kernel(int64_t n_elem, int npe,
                                                                     const int32_t *__restrict__ conn,
                                                                     const double *__restrict__ xyz,
                                                                     const int32_t *__restrict__ node_row,
                                                                     const double *__restrict__ params,
                                                                     const double *__restrict__ T,
                                                                     const double2 *__restrict__ scale, int64_t n,
                                                                     const double2 *__restrict__ X,
                                                                     double2 *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (e >= n_elem) return;
  double x[8][3];
  int32_t row[8];
  for (int k = 0; k < 8; k++) {
    const int32_t r = conn[e * npe + k];
    for (int j = 0; j < 3; j++) x[k][j] = xyz[(int64_t)r * 3 + j];
    row[k] = node_row[r];
  }
  double J[3][3], inv[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int k = 0; k < 8; k++) s += T[i * 8 + k] * x[k][j];
      J[i][j] = s;
    }
  jac_inverse(J, inv);
  double2 eps[6] = {};
  const double2 *Xf = X + (int64_t)f * n;
  for (int k = 0; k < 8; k++) {
    const double nx = inv[0][0] * T[k] + inv[0][1] * T[8 + k] + inv[0][2] * T[16 + k];
    const double ny = inv[1][0] * T[k] + inv[1][1] * T[8 + k] + inv[1][2] * T[16 + k];
    const double nz = inv[2][0] * T[k] + inv[2][1] * T[8 + k] + inv[2][2] * T[16 + k];
    const double2 u0 = Xf[row[k]], u1 = Xf[row[k] + 1], u2 = Xf[row[k] + 2];
    eps[0].x += nx * u0.x; eps[0].y += nx * u0.y;
    eps[1].x += ny * u1.x; eps[1].y += ny * u1.y;
    eps[2].x += nz * u2.x; eps[2].y += nz * u2.y;
    eps[3].x += ny * u0.x; eps[3].y += ny * u0.y;
    eps[3].x += nx * u1.x; eps[3].y += nx * u1.y;
    eps[4].x += nz * u1.x; eps[4].y += nz * u1.y;
    eps[4].x += ny * u2.x; eps[4].y += ny * u2.y;
    eps[5].x += nz * u0.x; eps[5].y += nz * u0.y;
    eps[5].x += nx * u2.x; eps[5].y += nx * u2.y;
  }
  // isotropic D (materialstructureisotrop.cpp:130-147): lambda, mu per element
  const double lam = params[e * 4], mu = params[e * 4 + 1];
  const double c11 = lam + 2.0 * mu;
  double2 sig[6];
  for (int c = 0; c < 2; c++) {
    const double e0 = c ? eps[0].y : eps[0].x, e1 = c ? eps[1].y : eps[1].x, e2 = c ? eps[2].y : eps[2].x;
    const double s0 = c11 * e0 + lam * e1 + lam * e2;
    const double s1 = lam * e0 + c11 * e1 + lam * e2;
    const double s2 = lam * e0 + lam * e1 + c11 * e2;
    if (c) { sig[0].y = s0; sig[1].y = s1; sig[2].y = s2; }
    else { sig[0].x = s0; sig[1].x = s1; sig[2].x = s2; }
  }
  for (int i = 3; i < 6; i++) sig[i] = double2{mu * eps[i].x, mu * eps[i].y};
  const double2 sc = scale[f];  // 1, or the loss factor's 1 + i eta (omega)
  double2 *o = out + ((int64_t)f * n_elem + e) * 6;
  for (int i = 0; i < 6; i++)
    o[i] = double2{sc.x * sig[i].x - sc.y * sig[i].y, sc.x * sig[i].y + sc.y * sig[i].x};
}

struct PostKind {
  const int32_t *adj_ptr = nullptr, *adj_elem = nullptr;
  const uint8_t *adj_local = nullptr;
  const double2 *res = nullptr;  // brick centroid stresses [f][e][6]
  int64_t n_elem = 0;
  int npe = 0, fluid = 0;
};
struct PostKinds {
  PostKind k[4];
};

// per node: the stresses of its bricks in ascending id, each scaled by
// 1 / (elements at the node) before it is added (elementstructurebrick.cpp:268-277);
// all other slots 0 (fluid_node_gradp_kernel adds grad p afterwards,
// elementfluid3d.cpp:395-404)
__global__ __launch_bounds__(kPostThreads) void node_results_kernel(int64_t n_nodes, const int8_t *__restrict__ slot,
                                                                   const int32_t *__restrict__ pos, PostKinds K,
                                                                   double2 *__restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (r >= n_nodes) return;
  double2 acc[9] = {};
  const int s = slot[r];
  if (s >= 0 && !K.k[s].fluid) {
    const PostKind &kb = K.k[s];
    const int32_t j0 = kb.adj_ptr[pos[r]], j1 = kb.adj_ptr[pos[r] + 1];
    const double w = 1.0 / (double)(j1 - j0);
    for (int32_t j = j0; j < j1; j++) {
      const double2 *g = kb.res + ((int64_t)f * kb.n_elem + kb.adj_elem[j]) * 6;
      for (int i = 0; i < 6; i++) {
        acc[i].x += g[i].x * w;
        acc[i].y += g[i].y * w;
      }
    }
  }
  double2 *o = out + ((int64_t)f * n_nodes + r) * 9;
  for (int i = 0; i < 9; i++) o[i] = acc[i];
}

// visco scale of D per frequency: E_vis = E (1 + i eta) (type 1) or
// E (1 + i omega eta) (type 2) (materialstructurevisco.cpp:283-285); D is
// linear in E.  Up to kScaleBatch frequencies per launch, passed by value.
constexpr int kScaleBatch = 64;
struct ScaleArgs {
  double w[kScaleBatch];
};
__global__ void visco_scale_kernel(ScaleArgs a, int nc, int has_visco, int visco_type, double eta,
                                   double2 *__restrict__ out) {
  const int f = threadIdx.x;
  if (f >= nc) return;
  const double im = has_visco ? (visco_type == 2 ? a.w[f] * eta : eta) : 0.0;
  out[f] = double2{1.0, im};
}

// Scratch of elp_node_results_noalloc: the scale factors of a chunk of
// frequencies and the brick centroid stresses of the chunk, which
// node_results_kernel averages per node.  Each device holds its own copy for
// the life of the process; host code uses the address hipGetSymbolAddress
// returns for the current device (kept in g_scratch_ptr), never the symbol.
constexpr int64_t kScratchBytes = 64ll << 20;
constexpr int64_t kScratchScaleBytes = kScaleBatch * (int64_t)sizeof(double2);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

}  // namespace

void post_scratch_init(int device) {
  ELP_REQUIRE(device >= 0 && device < kMaxDevices, ELP_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  ELP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the caller made `device` current
  g_scratch_ptr[device] = p;
}

int64_t post_scratch_bytes() { return kScratchBytes; }

// node -> (kind slot, entry of the kind's node list) and the derivative
// tables (fluid kinds at their nodes, bricks at the centroid), once per model
void post_prepare(Model &m) {
  Ctx *c = m.ctx;
  const int64_t N = m.n_nodes;
  std::vector<int8_t> slot(N, -1);
  std::vector<int32_t> pos(N, 0);
  for (size_t k = 0; k < m.kinds.size(); k++)
    for (size_t i = 0; i < m.kinds[k].nodes.size(); i++) {
      slot[m.kinds[k].nodes[i]] = (int8_t)k;
      pos[m.kinds[k].nodes[i]] = (int32_t)i;
    }
  m.d_post_slot.upload(c, slot);
  m.d_post_pos.upload(c, pos);
  // runs of whole nodes with at most kGradItems adjacency entries, one block
  // of fluid_item_gradp_kernel each (none if a node has more, or none, or the
  // solution of one frequency is too long for 32-bit offsets)
  for (auto &kb : m.kinds) {
    kb.grad_blk.clear();
    if (!is_fluid(kb.kind) || kb.nodes.empty()) continue;
    const int32_t nk = (int32_t)kb.nodes.size();
    std::vector<int32_t> b{0};
    bool fits = m.n * 16 < (int64_t)INT32_MAX;
    for (int32_t i = 0; i < nk && fits; i++) {
      fits = kb.adj_ptr[i + 1] - kb.adj_ptr[i] >= 1 && kb.adj_ptr[i + 1] - kb.adj_ptr[i] <= kGradItems;
      if (kb.adj_ptr[i + 1] - kb.adj_ptr[b.back()] > kGradItems || i - b.back() == kGradItems) b.push_back(i);
    }
    b.push_back(nk);
    if (fits) {
      kb.grad_blk = b;
      kb.d_grad_blk.upload(c, b);
    }
  }
  m.d_T8.upload(c, hex_nodal_derivs(8));
  m.d_T20.upload(c, hex_nodal_derivs(20));
  const std::vector<std::array<double, 3>> centroid(1, std::array<double, 3>{0.0, 0.0, 0.0});
  m.d_Tc.upload(c, hex_derivs_at(8, centroid));
}

namespace {

// The kernels of one chunk of nc frequencies: solutions dX [f][row], node
// results into dN [f][node][9] (may be null), brick stresses into res[k]
// [f][element][6] of each brick kind k.
void launch_post(Model &m, int nc, const double *w, const double2 *dX, double2 *dN, double2 *const *res,
                 double2 *dScale, hipStream_t s) {
  const int64_t N = m.n_nodes, n = m.n;
  ScaleArgs sa{};
  for (int f = 0; f < nc; f++) sa.w[f] = w[f];
  hipLaunchKernelGGL(visco_scale_kernel, dim3(1), dim3(kScaleBatch), 0, s, sa, nc, (int)m.has_visco, m.visco_type,
                     m.eta, dScale);
  ELP_HIP(hipGetLastError());
  PostKinds K;
  for (size_t k = 0; k < m.kinds.size(); k++) {
    KindBlock &kb = m.kinds[k];
    K.k[k] = PostKind{kb.d_adj_ptr.p, kb.d_adj_elem.p, kb.d_adj_local.p, res[k], kb.n_elem, kb.npe,
                      (int)is_fluid(kb.kind)};
    if (kb.n_elem == 0 || is_fluid(kb.kind)) continue;
    dim3 grid((unsigned)((kb.n_elem + kPostThreads - 1) / kPostThreads), (unsigned)nc);
    hipLaunchKernelGGL(brick_centroid_kernel, grid, dim3(kPostThreads), 0, s, kb.n_elem, kb.npe, kb.d_conn.p,
                       m.d_xyz.p, m.d_node_row.p, kb.d_params.p, m.d_Tc.p, dScale, n, dX, res[k]);
    ELP_HIP(hipGetLastError());
  }
  if (!dN) return;
  dim3 grid((unsigned)((N + kPostThreads - 1) / kPostThreads), (unsigned)nc);
  hipLaunchKernelGGL(node_results_kernel, grid, dim3(kPostThreads), 0, s, N, m.d_post_slot.p, m.d_post_pos.p, K, dN);
  ELP_HIP(hipGetLastError());
  for (auto &kb : m.kinds) {
    const int64_t nk = (int64_t)kb.nodes.size();
    if (!is_fluid(kb.kind) || nk == 0) continue;
    if (!kb.grad_blk.empty()) {
      const int32_t nr = (int32_t)kb.grad_blk.size() - 1;
      const dim3 gb((unsigned)nr);
      if (kb.kind == ELP_FLUID8)
        hipLaunchKernelGGL(fluid_item_gradp_kernel<8>, gb, dim3(kGradItems), 0, s, kb.d_grad_blk.p,
                           kb.d_nodes.p, kb.d_adj_ptr.p, kb.d_adj_elem.p, kb.d_adj_local.p, kb.d_conn.p, m.d_xyz.p,
                           m.d_node_row.p, m.d_T8.p, nc, n, N, dX, dN);
      else
        hipLaunchKernelGGL(fluid_item_gradp_kernel<20>, gb, dim3(kGradItems), 0, s, kb.d_grad_blk.p,
                           kb.d_nodes.p, kb.d_adj_ptr.p, kb.d_adj_elem.p, kb.d_adj_local.p, kb.d_conn.p, m.d_xyz.p,
                           m.d_node_row.p, m.d_T20.p, nc, n, N, dX, dN);
      ELP_HIP(hipGetLastError());
      continue;
    }
    dim3 g1((unsigned)((nk + kPostThreads - 1) / kPostThreads));
    if (kb.kind == ELP_FLUID8)
      hipLaunchKernelGGL(fluid_node_gradp_kernel<8>, g1, dim3(kPostThreads), 0, s, nk, kb.d_nodes.p, kb.d_adj_ptr.p,
                         kb.d_adj_elem.p, kb.d_adj_local.p, kb.d_conn.p, m.d_xyz.p, m.d_node_row.p, m.d_T8.p, nc, n,
                         N, dX, dN);
    else
      hipLaunchKernelGGL(fluid_node_gradp_kernel<20>, g1, dim3(kPostThreads), 0, s, nk, kb.d_nodes.p,
                         kb.d_adj_ptr.p, kb.d_adj_elem.p, kb.d_adj_local.p, kb.d_conn.p, m.d_xyz.p, m.d_node_row.p,
                         m.d_T20.p, nc, n, N, dX, dN);
    ELP_HIP(hipGetLastError());
  }
}

int64_t brick_kinds_of(const Model &m) {
  int64_t b = 0;
  for (auto &kb : m.kinds) b += !is_fluid(kb.kind);
  return b;
}

}  // namespace

void node_results(Model &m, int64_t nf, const double *freqs, const double *X, double *cells, double *nodes,
                  double *kernel_ms) {
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  Ctx *c = m.ctx;
  hipStream_t s = c->stream;
  const int64_t N = m.n_nodes, n = m.n;
  // frequency chunk: solutions, node results and brick stresses within
  // about 4 GiB (the fluid gradient operators are reused across a chunk)
  int64_t per_f = n * 16 + (nodes ? N * 9 * 16 : 0);
  for (auto &kb : m.kinds)
    if (!is_fluid(kb.kind)) per_f += kb.n_elem * 6 * 16;
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>({nf, (int64_t)(4LL << 30) / per_f, kScaleBatch}));
  DArr<double2> dX, dN, dScale;
  std::vector<DArr<double2>> dres(m.kinds.size());
  dX.alloc(c, chunk * n);
  if (nodes) dN.alloc(c, chunk * N * 9);
  dScale.alloc(c, kScaleBatch);
  std::vector<double2 *> res(m.kinds.size(), nullptr);
  for (size_t k = 0; k < m.kinds.size(); k++)
    if (!is_fluid(m.kinds[k].kind)) {
      dres[k].alloc(c, chunk * m.kinds[k].n_elem * 6);
      res[k] = dres[k].p;
    }
  const int64_t brick_cols = brick_element_count(m);
  ELP_REQUIRE(!cells || brick_kinds_of(m) <= 1, ELP_EINVAL,
              "element stresses of a model mixing Brick8 and Brick20 elements are not supported");
  std::unique_ptr<HipEvent> ev0, ev1;
  if (kernel_ms) {
    *kernel_ms = 0.0;
    ev0.reset(new HipEvent);
    ev1.reset(new HipEvent);
  }
  std::vector<double> w(kScaleBatch);
  for (int64_t f0 = 0; f0 < nf; f0 += chunk) {
    const int nc = (int)std::min(chunk, nf - f0);
    ELP_HIP(hipMemcpyAsync(dX.p, X + 2 * f0 * n, sizeof(double2) * nc * n, hipMemcpyHostToDevice, s));
    for (int f = 0; f < nc; f++) w[f] = 2.0 * M_PI * freqs[f0 + f];
    if (kernel_ms) ELP_HIP(hipEventRecord(ev0->e, s));
    launch_post(m, nc, w.data(), dX.p, nodes ? dN.p : nullptr, res.data(), dScale.p, s);
    if (kernel_ms) ELP_HIP(hipEventRecord(ev1->e, s));
    if (nodes)
      ELP_HIP(hipMemcpyAsync(nodes + 2 * f0 * N * 9, dN.p, sizeof(double2) * nc * N * 9, hipMemcpyDeviceToHost, s));
    if (cells && brick_cols) {
      // [f][brick element][6]
      for (int f = 0; f < nc; f++) {
        int64_t col = 0;
        for (size_t k = 0; k < m.kinds.size(); k++) {
          KindBlock &kb = m.kinds[k];
          if (is_fluid(kb.kind) || kb.n_elem == 0) continue;
          ELP_HIP(hipMemcpyAsync(cells + 2 * 6 * ((f0 + f) * brick_cols + col), dres[k].p + (int64_t)f * kb.n_elem * 6,
                                 sizeof(double2) * kb.n_elem * 6, hipMemcpyDeviceToHost, s));
          col += kb.n_elem;
        }
      }
    }
    ELP_HIP(hipStreamSynchronize(s));
    if (kernel_ms) {
      float ms = 0.f;
      ELP_HIP(hipEventElapsedTime(&ms, ev0->e, ev1->e));
      *kernel_ms += ms;
    }
  }
}

void node_results_noalloc(Model &m, int64_t nf, const double *freqs, const double *X_dev, double *nodes_dev) {
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  Ctx *c = m.ctx;
  hipStream_t s = c->stream;
  const int64_t N = m.n_nodes, n = m.n;
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = static_cast<unsigned char *>(g_scratch_ptr[c->device]);
  }
  ELP_REQUIRE(base, ELP_ESTATE, "scratch of the device not set up (create the context first)");
  // per frequency: the centroid stresses of every brick, each kind's block
  // starting on a 256-byte boundary
  int64_t per_f = 0;
  for (auto &kb : m.kinds)
    if (!is_fluid(kb.kind)) per_f += (kb.n_elem * 6 * 16 + 255) / 256 * 256;
  const int64_t room = kScratchBytes - kScratchScaleBytes;
  ELP_REQUIRE(per_f <= room, ELP_EINVAL,
              "the model's " + std::to_string(brick_element_count(m)) + " brick elements need " +
                  std::to_string(per_f) + " bytes of scratch per frequency; the allocation-free node results hold " +
                  std::to_string(room) + " (at most " + std::to_string(room / 96) + " bricks)");
  const int64_t chunk = std::min<int64_t>({nf, per_f ? room / per_f : (int64_t)kScaleBatch, kScaleBatch});
  double2 *dScale = reinterpret_cast<double2 *>(base);
  std::vector<double2 *> res(m.kinds.size(), nullptr);
  double w[kScaleBatch];
  for (int64_t f0 = 0; f0 < nf; f0 += chunk) {
    const int nc = (int)std::min(chunk, nf - f0);
    unsigned char *q = base + kScratchScaleBytes;
    for (size_t k = 0; k < m.kinds.size(); k++) {
      const KindBlock &kb = m.kinds[k];
      if (is_fluid(kb.kind)) continue;
      res[k] = reinterpret_cast<double2 *>(q);
      q += (nc * kb.n_elem * 6 * 16 + 255) / 256 * 256;
    }
    for (int f = 0; f < nc; f++) w[f] = 2.0 * M_PI * freqs[f0 + f];
    launch_post(m, nc, w, reinterpret_cast<const double2 *>(X_dev) + f0 * n,
                reinterpret_cast<double2 *>(nodes_dev) + f0 * N * 9, res.data(), dScale, s);
  }
  ELP_HIP(hipStreamSynchronize(s));
}

int64_t brick_element_count(const Model &m) {
  int64_t b = 0;
  for (auto &kb : m.kinds)
    if (!is_fluid(kb.kind)) b += kb.n_elem;
  return b;
}

}  // namespace elp
