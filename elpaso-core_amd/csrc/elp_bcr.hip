// Block cyclic reduction of the structure block (see elp_bcr.h).
//
// Storage: every block is m x m, column-major (element (a, b) at b m + a),
// complex; right-hand sides of a solve are N blocks of m x B (slot c of
// entry a at c m + a), reduced in place: the blocks of level l sit at
// original indices 2^l p + 2^l - 1 (position p of the level), so a level's
// kept (odd p) and eliminated (even p) blocks are strided views of one array.
#include "elp_bcr.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <tuple>

namespace elp {

#define ELP_RB(call)                                                                                    \
  do {                                                                                                  \
    rocblas_status st_ = (call);                                                                        \
    if (st_ != rocblas_status_success)                                                                  \
      throw ::elp::Error(ELP_EHIP, std::string(#call) + ": rocBLAS status " + std::to_string((int)st_)); \
  } while (0)

template <typename T> struct CT;
template <> struct CT<float> {
  using v = float2;
  using rb = rocblas_float_complex;
};
template <> struct CT<double> {
  using v = double2;
  using rb = rocblas_double_complex;
};

struct BcrPlan {
  int64_t ns = 0;  // free structure rows
  int m = 0, N = 0, axis = 0;
  DArr<int32_t> rows_model, rows_solver, pos;
  // separate handles for solves (h) and factorizations (hf): a factorization
  // may run on another thread and stream while solves go on
  rocblas_handle h = nullptr, hf = nullptr;
  DArr<unsigned char> X, Tw;  // solve work: N blocks and ceil(N/2) blocks of m x B
  // factorization work (one factorization at a time), kept between
  // factorizations so that none frees device memory (hipFree waits for the
  // whole device): D and C of a level, ping-pong, pivots
  DArr<unsigned char> fD[2], fC[2];
  DArr<int> ipiv, info, bad;
  ~BcrPlan() {
    if (h) (void)rocblas_destroy_handle(h);
    if (hf) (void)rocblas_destroy_handle(hf);
  }
};

template <typename T>
struct BcrFactor {
  using V = typename CT<T>::v;
  struct Level {
    int N = 0, E = 0, K = 0, K1 = 0;  // blocks, eliminated, kept, kept with a right neighbour
    DArr<V> Dinv, W, Vv;
  };
  std::vector<Level> lv;
};

static void free_plan(BcrPlan *p) { delete p; }
template <typename T>
static void free_factor(BcrFactor<T> *f) {
  delete f;
}

int bcr_block_size(const BcrPlan &p) { return p.m; }
int bcr_blocks(const BcrPlan &p) { return p.N; }
int64_t bcr_rows(const BcrPlan &p) { return p.ns; }
int64_t bcr_factor_bytes(const BcrPlan &p, int elem_bytes) {
  // kept after factoring: Dinv, W, V of every level, about 1.5 N blocks at
  // level 0 and half as many per level above; peak while factoring: D, C
  // of a level besides
  const int64_t m2 = (int64_t)p.m * p.m * elem_bytes;
  return (int64_t)(3.0 * p.N + 4) * m2 + 2 * (int64_t)p.N * m2;
}

template <typename T>
std::unique_ptr<BcrFactor<T>, void (*)(BcrFactor<T> *)> bcr_new_factor() {
  return std::unique_ptr<BcrFactor<T>, void (*)(BcrFactor<T> *)>(new BcrFactor<T>(), free_factor<T>);
}
template std::unique_ptr<BcrFactor<float>, void (*)(BcrFactor<float> *)> bcr_new_factor<float>();
template std::unique_ptr<BcrFactor<double>, void (*)(BcrFactor<double> *)> bcr_new_factor<double>();

// ---------------------------------------------------------------------------
// the plan: ordering, block size, maps
// ---------------------------------------------------------------------------
std::unique_ptr<BcrPlan, void (*)(BcrPlan *)> bcr_plan(const Model &m, const std::vector<int32_t> *iperm) {
  std::unique_ptr<BcrPlan, void (*)(BcrPlan *)> none(nullptr, free_plan);
  // free structure rows and their nodes
  std::vector<int32_t> row_node(m.n, -1);
  for (int64_t r = 0; r < m.n_nodes; r++)
    for (int d = 0; d < m.node_nd[r]; d++)
      if (m.node_row[r] >= 0) row_node[m.node_row[r] + d] = (int32_t)r;
  std::vector<int32_t> nodes;
  for (int64_t r = 0; r < m.n_nodes; r++) {
    if (m.node_row[r] < 0) continue;
    bool any = false;
    for (int d = 0; d < m.node_nd[r]; d++) any |= m.row_field[m.node_row[r] + d] == 0;
    if (any) nodes.push_back((int32_t)r);
  }
  if (nodes.empty()) return none;
  auto coord = [&](int32_t rank, int d) { return m.xyz[(int64_t)m.rank_to_input[rank] * 3 + d]; };
  int best_axis = 0;
  int64_t best_b = INT64_MAX;
  std::vector<int32_t> best_pos;
  std::vector<int32_t> pos(m.n);
  for (int a = 0; a < 3; a++) {
    std::vector<int32_t> ord(nodes);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t u, int32_t v) {
      return std::make_tuple(coord(u, a), coord(u, (a + 1) % 3), coord(u, (a + 2) % 3)) <
             std::make_tuple(coord(v, a), coord(v, (a + 1) % 3), coord(v, (a + 2) % 3));
    });
    std::fill(pos.begin(), pos.end(), -1);
    int32_t q = 0;
    for (int32_t rank : ord)
      for (int d = 0; d < m.node_nd[rank]; d++) {
        const int32_t row = m.node_row[rank] + d;
        if (m.row_field[row] == 0) pos[row] = q++;
      }
    int64_t b = 0;
    for (int64_t r = 0; r < m.n; r++) {
      if (pos[r] < 0) continue;
      for (int32_t p = m.rowptr[r]; p < m.rowptr[r + 1]; p++) {
        const int32_t c = m.colind[p];
        if (pos[c] >= 0) b = std::max<int64_t>(b, std::abs((int64_t)pos[r] - pos[c]));
      }
    }
    if (b < best_b) {
      best_b = b;
      best_axis = a;
      best_pos = pos;
    }
  }
  auto P = std::unique_ptr<BcrPlan, void (*)(BcrPlan *)>(new BcrPlan(), free_plan);
  int64_t ns = 0;
  for (int32_t v : best_pos) ns += v >= 0;
  P->ns = ns;
  P->axis = best_axis;
  P->m = (int)std::max<int64_t>(1, best_b);
  ELP_REQUIRE(P->m <= 8192, ELP_EINVAL,
              "structure block too wide for the direct coupled preconditioner (half-bandwidth " +
                  std::to_string(best_b) + ")");
  P->N = (int)((ns + P->m - 1) / P->m);
  std::vector<int32_t> rm((size_t)P->N * P->m, -1), rs((size_t)P->N * P->m, -1);
  for (int64_t r = 0; r < m.n; r++)
    if (best_pos[r] >= 0) {
      rm[best_pos[r]] = (int32_t)r;
      rs[best_pos[r]] = iperm ? (*iperm)[r] : (int32_t)r;
    }
  P->rows_model.upload(m.ctx, rm);
  P->rows_solver.upload(m.ctx, rs);
  P->pos.upload(m.ctx, best_pos);
  ELP_RB(rocblas_create_handle(&P->h));
  ELP_RB(rocblas_create_handle(&P->hf));
  return P;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// level-0 blocks: D_i (rows and columns in block i) and C_i (rows in block
// i, columns in block i - 1) of A = ks K + ms M; padding rows identity
template <typename V>
__global__ void bcr_extract_kernel(int64_t total, int m, const int32_t *__restrict__ rows_model,
                                   const int32_t *__restrict__ pos, const int32_t *__restrict__ rowptr,
                                   const int32_t *__restrict__ colind, const double *__restrict__ K,
                                   const double *__restrict__ M, double2 ks, double2 ms, V *__restrict__ D,
                                   V *__restrict__ C, int *__restrict__ bad) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int64_t i = idx / m, a = idx % m, m2 = (int64_t)m * m;
  const int32_t r = rows_model[idx];
  if (r < 0) {
    D[i * m2 + a * m + a] = V{1, 0};
    return;
  }
  for (int32_t p = rowptr[r]; p < rowptr[r + 1]; p++) {
    const int32_t q = pos[colind[p]];
    if (q < 0) continue;
    const int64_t j = q / m, b = q % m;
    V v;
    v.x = ks.x * K[p] + ms.x * M[p];
    v.y = ks.y * K[p] + ms.y * M[p];
    if (j == i) D[i * m2 + b * m + a] = v;
    else if (j == i - 1) C[i * m2 + b * m + a] = v;
    else if (j != i + 1) *bad = 1;
  }
}

// dst block k = src block k (count blocks of bs elements, strides in elements)
template <typename V>
__global__ void bcr_copy_blocks_kernel(int64_t count, int64_t bs, const V *__restrict__ src, int64_t sstride,
                                       V *__restrict__ dst, int64_t dstride) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count * bs) return;
  const int64_t k = i / bs, e = i % bs;
  dst[k * dstride + e] = src[k * sstride + e];
}

// X[blk][c][a] = r[row][c] (solver order, B columns per row, converted to
// the factor's precision), 0 on padding
template <typename VX, typename VR>
__global__ void bcr_gather_kernel(int64_t total, int m, int B, const int32_t *__restrict__ rows,
                                  const VR *__restrict__ r, VX *__restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total * B) return;
  const int64_t idx = i / B;
  const int c = (int)(i % B);
  const int64_t blk = idx / m, a = idx % m;
  const int32_t row = rows[idx];
  VX v = {0, 0};
  if (row >= 0) {
    const VR t = r[(int64_t)row * B + c];
    v.x = t.x;
    v.y = t.y;
  }
  X[(blk * B + c) * m + a] = v;
}

template <typename VX, typename VR>
__global__ void bcr_scatter_kernel(int64_t total, int m, int B, const int32_t *__restrict__ rows,
                                   const VX *__restrict__ X, VR *__restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total * B) return;
  const int64_t idx = i / B;
  const int c = (int)(i % B);
  const int32_t row = rows[idx];
  if (row < 0) return;
  const int64_t blk = idx / m, a = idx % m;
  const VX t = X[(blk * B + c) * m + a];
  VR v;
  v.x = t.x;
  v.y = t.y;
  z[(int64_t)row * B + c] = v;
}

static int nb256(int64_t n) { return (int)((n + 255) / 256); }

// ---------------------------------------------------------------------------
// rocBLAS / rocSOLVER wrappers (column-major, strided batched)
// ---------------------------------------------------------------------------
static rocblas_operation op(bool t) { return t ? rocblas_operation_transpose : rocblas_operation_none; }

static void gemm_sb(rocblas_handle h, bool ta, bool tb, int M, int N, int K, float alpha, const float2 *A,
                    int64_t sa, const float2 *B, int ldb, int64_t sb, float beta, float2 *C, int ldc, int64_t sc,
                    int batch) {
  if (batch <= 0) return;
  const rocblas_float_complex al = {alpha, 0.f}, be = {beta, 0.f};
  ELP_RB(rocblas_cgemm_strided_batched(h, op(ta), op(tb), M, N, K, &al, (const rocblas_float_complex *)A, M, sa,
                                       (const rocblas_float_complex *)B, ldb, sb, &be, (rocblas_float_complex *)C,
                                       ldc, sc, batch));
}
static void gemm_sb(rocblas_handle h, bool ta, bool tb, int M, int N, int K, double alpha, const double2 *A,
                    int64_t sa, const double2 *B, int ldb, int64_t sb, double beta, double2 *C, int ldc, int64_t sc,
                    int batch) {
  if (batch <= 0) return;
  const rocblas_double_complex al = {alpha, 0.0}, be = {beta, 0.0};
  ELP_RB(rocblas_zgemm_strided_batched(h, op(ta), op(tb), M, N, K, &al, (const rocblas_double_complex *)A, M, sa,
                                       (const rocblas_double_complex *)B, ldb, sb, &be, (rocblas_double_complex *)C,
                                       ldc, sc, batch));
}
static void ensure_bytes(DArr<unsigned char> &a, Ctx *c, int64_t bytes) {
  if ((int64_t)a.n < bytes) a.alloc(c, bytes);
}
template <typename V>
static void ensure(DArr<V> &a, Ctx *c, int64_t count) {
  if ((int64_t)a.n < count) a.alloc(c, count);
}

static void inverse_sb(rocblas_handle h, int m, float2 *A, int batch, rocblas_int *ipiv, rocblas_int *info) {
  const int64_t m2 = (int64_t)m * m;
  ELP_RB(rocsolver_cgetrf_strided_batched(h, m, m, (rocblas_float_complex *)A, m, m2, ipiv, m, info, batch));
  ELP_RB(rocsolver_cgetri_strided_batched(h, m, (rocblas_float_complex *)A, m, m2, ipiv, m, info, batch));
}
static void inverse_sb(rocblas_handle h, int m, double2 *A, int batch, rocblas_int *ipiv, rocblas_int *info) {
  const int64_t m2 = (int64_t)m * m;
  ELP_RB(rocsolver_zgetrf_strided_batched(h, m, m, (rocblas_double_complex *)A, m, m2, ipiv, m, info, batch));
  ELP_RB(rocsolver_zgetri_strided_batched(h, m, (rocblas_double_complex *)A, m, m2, ipiv, m, info, batch));
}

template <typename V>
static void copy_blocks(int64_t count, int64_t bs, const V *src, int64_t sstride, V *dst, int64_t dstride,
                        hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(bcr_copy_blocks_kernel<V>, dim3(nb256(count * bs)), dim3(256), 0, s, count, bs, src, sstride,
                     dst, dstride);
  ELP_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// factorization
// ---------------------------------------------------------------------------
template <typename T>
void bcr_factor(const Model &m, BcrPlan &p, std::complex<double> ks, std::complex<double> ms, BcrFactor<T> &f,
                hipStream_t s) {
  using V = typename CT<T>::v;
  const int mb = p.m;
  const int64_t m2 = (int64_t)mb * mb;
  Ctx *cx = m.ctx;
  ELP_RB(rocblas_set_stream(p.hf, s));
  for (int k = 0; k < 2; k++) {
    ensure_bytes(p.fD[k], cx, (int64_t)p.N * m2 * (int64_t)sizeof(V));
    ensure_bytes(p.fC[k], cx, (int64_t)p.N * m2 * (int64_t)sizeof(V));
  }
  const int E0 = (p.N + 1) / 2;
  ensure(p.ipiv, cx, (int64_t)E0 * mb);
  ensure(p.info, cx, E0);
  ensure(p.bad, cx, 1);
  V *D = reinterpret_cast<V *>(p.fD[0].p), *C = reinterpret_cast<V *>(p.fC[0].p);
  V *Dn = reinterpret_cast<V *>(p.fD[1].p), *Cn = reinterpret_cast<V *>(p.fC[1].p);
  ELP_HIP(hipMemsetAsync(D, 0, sizeof(V) * p.N * m2, s));
  ELP_HIP(hipMemsetAsync(C, 0, sizeof(V) * p.N * m2, s));
  ELP_HIP(hipMemsetAsync(p.bad.p, 0, sizeof(int), s));
  const int64_t total = (int64_t)p.N * mb;
  hipLaunchKernelGGL(bcr_extract_kernel<V>, dim3(nb256(total)), dim3(256), 0, s, total, mb, p.rows_model.p, p.pos.p,
                     m.d_rowptr.p, m.d_colind.p, m.d_K.p, m.d_M.p, double2{ks.real(), ks.imag()},
                     double2{ms.real(), ms.imag()}, D, C, p.bad.p);
  ELP_HIP(hipGetLastError());
  int hbad = 0;
  ELP_HIP(hipMemcpyAsync(&hbad, p.bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  ELP_HIP(hipStreamSynchronize(s));
  ELP_REQUIRE(!hbad, ELP_EINVAL, "structure block is not block tridiagonal in its ordering");
  // the level structure depends on N only: a refactorization reuses the
  // level arrays
  int N = p.N;
  size_t l = 0;
  while (true) {
    if (f.lv.size() <= l) f.lv.emplace_back();
    typename BcrFactor<T>::Level &L = f.lv[l];
    L.N = N;
    L.E = (N + 1) / 2;
    L.K = N / 2;
    L.K1 = (N % 2) ? L.K : L.K - 1;
    const int K = L.K, K1 = L.K1;
    // inverses of the eliminated (even-position) blocks
    ensure(L.Dinv, cx, (int64_t)L.E * m2);
    copy_blocks<V>(L.E, m2, D, 2 * m2, L.Dinv.p, m2, s);
    inverse_sb(p.hf, mb, L.Dinv.p, L.E, p.ipiv.p, p.info.p);
    if (K == 0) {
      f.lv.resize(l + 1);
      break;
    }
    ensure(L.W, cx, (int64_t)K * m2);
    ensure(L.Vv, cx, (int64_t)K * m2);
    // W_k = C_k D_{k-1}^-1 ; V_k = C_{k+1}^T D_{k+1}^-1
    gemm_sb(p.hf, false, false, mb, mb, mb, (T)1, C + m2, 2 * m2, L.Dinv.p, mb, m2, (T)0, L.W.p, mb, m2, K);
    gemm_sb(p.hf, true, false, mb, mb, mb, (T)1, C + 2 * m2, 2 * m2, L.Dinv.p + m2, mb, m2, (T)0, L.Vv.p, mb, m2,
            K1);
    if (K1 < K) ELP_HIP(hipMemsetAsync(L.Vv.p + (size_t)(K - 1) * m2, 0, sizeof(V) * m2, s));
    // D'_k = D_k - W_k C_k^T - V_k C_{k+1}
    copy_blocks<V>(K, m2, D + m2, 2 * m2, Dn, m2, s);
    gemm_sb(p.hf, false, true, mb, mb, mb, (T)-1, L.W.p, m2, C + m2, mb, 2 * m2, (T)1, Dn, mb, m2, K);
    gemm_sb(p.hf, false, false, mb, mb, mb, (T)-1, L.Vv.p, m2, C + 2 * m2, mb, 2 * m2, (T)1, Dn, mb, m2, K1);
    // C'_k = -W_k C_{k-1} (coupling of kept k to the kept block before it)
    ELP_HIP(hipMemsetAsync(Cn, 0, sizeof(V) * m2, s));
    gemm_sb(p.hf, false, false, mb, mb, mb, (T)-1, L.W.p + m2, m2, C + 2 * m2, mb, 2 * m2, (T)0, Cn + m2, mb, m2,
            K - 1);
    std::swap(D, Dn);
    std::swap(C, Cn);
    N = K;
    l++;
  }
  ELP_HIP(hipStreamSynchronize(s));
}

template void bcr_factor<float>(const Model &, BcrPlan &, std::complex<double>, std::complex<double>,
                                BcrFactor<float> &, hipStream_t);
template void bcr_factor<double>(const Model &, BcrPlan &, std::complex<double>, std::complex<double>,
                                 BcrFactor<double> &, hipStream_t);

// ---------------------------------------------------------------------------
// solve
// ---------------------------------------------------------------------------
template <typename T, typename VR>
void bcr_apply(BcrPlan &p, const BcrFactor<T> &f, int B, const VR *r, VR *z, hipStream_t s, bool model_order) {
  const int32_t *rows = model_order ? p.rows_model.p : p.rows_solver.p;
  using V = typename CT<T>::v;
  const int mb = p.m;
  const int64_t m2 = (int64_t)mb * mb, blk = (int64_t)mb * B;
  const int64_t xbytes = (int64_t)p.N * blk * (int64_t)sizeof(V), tbytes = (int64_t)((p.N + 1) / 2) * blk * sizeof(V);
  if ((int64_t)p.X.n < xbytes) p.X.alloc(p.rows_model.ctx, xbytes);
  if ((int64_t)p.Tw.n < tbytes) p.Tw.alloc(p.rows_model.ctx, tbytes);
  V *X = reinterpret_cast<V *>(p.X.p);
  V *Tw = reinterpret_cast<V *>(p.Tw.p);
  ELP_RB(rocblas_set_stream(p.h, s));
  const int64_t total = (int64_t)p.N * mb;
  hipLaunchKernelGGL((bcr_gather_kernel<V, VR>), dim3(nb256(total * B)), dim3(256), 0, s, total, mb, B,
                     rows, r, X);
  ELP_HIP(hipGetLastError());
  const int L = (int)f.lv.size();
  // reduction: b_k -= W_k b_{k-1} + V_k b_{k+1} on the kept blocks of each level
  for (int l = 0; l < L - 1; l++) {
    const auto &Lv = f.lv[l];
    const int64_t S = 1LL << l, off = S - 1;
    gemm_sb(p.h, false, false, mb, B, mb, (T)-1, Lv.W.p, m2, X + off * blk, mb, 2 * S * blk, (T)1,
            X + (off + S) * blk, mb, 2 * S * blk, Lv.K);
    gemm_sb(p.h, false, false, mb, B, mb, (T)-1, Lv.Vv.p, m2, X + (off + 2 * S) * blk, mb, 2 * S * blk, (T)1,
            X + (off + S) * blk, mb, 2 * S * blk, Lv.K1);
  }
  // back-substitution: x_e = D_e^-1 b_e - V_{e-1}^T x_{e-1} - W_{e+1}^T x_{e+1}
  for (int l = L - 1; l >= 0; l--) {
    const auto &Lv = f.lv[l];
    const int64_t S = 1LL << l, off = S - 1;
    gemm_sb(p.h, false, false, mb, B, mb, (T)1, Lv.Dinv.p, m2, X + off * blk, mb, 2 * S * blk, (T)0, Tw, mb, blk,
            Lv.E);
    if (Lv.K > 0) {
      gemm_sb(p.h, true, false, mb, B, mb, (T)-1, Lv.Vv.p, m2, X + (off + S) * blk, mb, 2 * S * blk, (T)1, Tw + blk,
              mb, blk, Lv.E - 1);
      gemm_sb(p.h, true, false, mb, B, mb, (T)-1, Lv.W.p, m2, X + (off + S) * blk, mb, 2 * S * blk, (T)1, Tw, mb,
              blk, Lv.K);
    }
    copy_blocks<V>(Lv.E, blk, Tw, blk, X + off * blk, 2 * S * blk, s);
  }
  hipLaunchKernelGGL((bcr_scatter_kernel<V, VR>), dim3(nb256(total * B)), dim3(256), 0, s, total, mb, B,
                     rows, X, z);
  ELP_HIP(hipGetLastError());
}

template void bcr_apply<float, float2>(BcrPlan &, const BcrFactor<float> &, int, const float2 *, float2 *,
                                       hipStream_t, bool);
template void bcr_apply<double, float2>(BcrPlan &, const BcrFactor<double> &, int, const float2 *, float2 *,
                                        hipStream_t, bool);
template void bcr_apply<double, double2>(BcrPlan &, const BcrFactor<double> &, int, const double2 *, double2 *,
                                         hipStream_t, bool);

}  // namespace elp
