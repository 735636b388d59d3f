// Harmonic sweep solver on one MI355X.
//
// The system of one frequency w (elPaSo's K_dyn(w), restructured so the
// global matrices are assembled once):
//   fluid rows      A = K/w^2 - M + i/(w Z) C_G       (K, M already / rho)
//   structure rows  A = (1 + i eta g + i w a) K + (-w^2 + i w b) M
//   plus the FSI block S; Dirichlet rows are identity rows (their columns are
//   zeroed in the solver's copies: prescribed values are zero).
// These operators are complex symmetric, so the Krylov method is COCG
// (conjugate orthogonal CG: CG with the unconjugated bilinear form), whose
// short recurrence keeps the memory per frequency at a handful of vectors.
//
// Vectors of B frequency slots are stored [row][B] (frequency fastest): one
// gather of a column fetches the B values of that column in one contiguous
// line, and K, M and the column indices are read once for all slots (the
// fused multi-frequency operator; for B = 32 row_op_kernel, one wavefront per
// row).  Mixed precision: the Krylov iteration and the preconditioner run in
// complex FP32 on FP32 copies of K and M; an outer FP64 iterative refinement
// recomputes the true residual with the FP64 matrices until ||r||/||b|| <=
// rtol.  A sweep runs on B slots that are refilled as frequencies finish
// (solve_slots).
//
// Preconditioners: for fluid box models the exact fast-diagonalization solve
// of the Hex8 box operator (elp_fdm), used directly for Hex8 models and as the
// coarse level of a symmetric two-level cycle (l1-Jacobi pre- and post-sweep)
// for Hex20, with the shifted operator P(w) = K/w^2 - (1 - i beta) M + ...
// (beta = 0 where the box solve is exact, see auto_shift);
// for a Hex8 fluid box coupled to a structure the block-diagonal
// preconditioner of the exact box solve (fluid rows) and the block cyclic
// reduction of the structure block (elp_bcr, factored per window of
// frequencies); otherwise a rediscretized Hex8 multigrid hierarchy
// (trilinear transfer, l1-Jacobi smoothing) or l1-Jacobi alone.  Every cycle is a fixed linear
// operator and complex symmetric (R = P^T, symmetric smoothing), as COCG
// requires.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>

#include "elp_bcr.h"
#include "elp_fdm.h"
#include "elp_internal.h"
#include "elp_solver.h"

namespace elp {

template <typename T> struct VT;
template <> struct VT<float> { using t = float2; };
template <> struct VT<double> { using t = double2; };

struct OpCoef {
  double2 kf, mf, ks, ms, g;       // complex coefficients
  double akf, amf, aks, ams, ag;   // magnitudes (l1 diagonal bound)
};

template <typename V>
__device__ inline V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

template <typename T>
struct OpArgs {
  int64_t n;
  const int32_t *rowptr, *colind;
  const T *K, *M;
  const int32_t *xrp, *xcol;
  const uint8_t *xtype;
  const T *xval;
  const uint8_t *field;
  const double *rK, *rM, *rG;
};

// Fused multi-frequency operator.  Each row is handled by a group of
// G = max(B, 8) lanes; lane l works on frequency f = l % B.  The row is walked
// in chunks of CH nonzeros: the group loads the chunk's column indices and K,
// M values with one coalesced load per lane into LDS, then every lane gathers
// x[col][f] for the chunk's entries it owns (entries j = sub, sub + S, ...,
// S = G / B) with independent loads, so the gathers of a chunk are in flight
// together.  Partial sums of the S subsets are combined with a fixed xor tree:
// the result is deterministic.
//   MODE 0: y = A x
//   MODE 1: y = b - A x
//   MODE 2: y = x + (b - A x) / D      (one l1-Jacobi sweep)
//   MODE 3: y = b / D                  (first sweep from a zero guess)
// Workgroups are dealt round-robin to the 8 XCDs of the MI355X, each with its
// own L2.  Launches use a multiple of 8 workgroups and remap them so that each
// XCD walks one contiguous range of rows: neighbouring rows share most of
// their columns, so the gathers of x hit in that XCD's L2.
constexpr int kXcds = 8;
// block `local` of a launch (or of a segment of a grouped launch) of
// `nblocks` blocks, a multiple of 8, remapped so that each XCD walks one
// contiguous range
__device__ inline int xcd_map(int local, int nblocks) {
  const int per = nblocks / kXcds;
  return (local % kXcds) * per + local / kXcds;
}
__device__ inline int xcd_block() { return xcd_map(blockIdx.x, gridDim.x); }
static int xcd_grid(int64_t blocks) { return (int)((blocks + kXcds - 1) / kXcds * kXcds); }

template <int B> struct OpShape {
  static constexpr int G = B < 8 ? 8 : B;    // lanes per row
  static constexpr int S = G / B;            // nonzero subsets per frequency
  static constexpr int U = B < 8 ? 4 : 1;    // chunk = U entries per lane
  static constexpr int CH = G * U;           // nonzeros per chunk
  static constexpr int RPB = 256 / G;        // rows per block
};

// (A x)[r][f] from a row's K, M and extra sums and the frequency's
// coefficients (fl: a fluid row)
template <typename T>
__device__ inline typename VT<T>::t op_value(const OpCoef &cf, bool fl, typename VT<T>::t accK,
                                             typename VT<T>::t accM, typename VT<T>::t accX) {
  using V = typename VT<T>::t;
  const V ck = fl ? V{(T)cf.kf.x, (T)cf.kf.y} : V{(T)cf.ks.x, (T)cf.ks.y};
  const V cm = fl ? V{(T)cf.mf.x, (T)cf.mf.y} : V{(T)cf.ms.x, (T)cf.ms.y};
  V ax = cmul(ck, accK), bx = cmul(cm, accM);
  V res;
  res.x = ax.x + bx.x + accX.x;
  res.y = ax.y + bx.y + accX.y;
  return res;
}

// FP64 residual of a few frozen columns on a narrow block of CB columns
// (MODE 4): x holds their x64 columns gathered, xn[r][c] = x64[r][map[c]],
// and r64[r][map[c]] = b64[r][map[c]] - (A x)[r][c] is stored for the nfz
// frozen columns c < nfz (the padding columns c >= nfz store nothing).
// b64 and r64 are [row][ld].
struct NarrowOut {
  const int *map;
  int nfz, ld;
  const double2 *b64;
  double2 *r64;
};
__device__ inline void narrow_store(const NarrowOut &nw, int64_t r, int c, double2 v) {
  if (c >= nw.nfz) return;
  const int64_t o = r * nw.ld + nw.map[c];
  const double2 bv = nw.b64[o];
  nw.r64[o] = {bv.x - v.x, bv.y - v.y};
}

// Row epilogue shared by the operator kernels: combine the K, M and extra
// sums with the frequency's coefficients and apply MODE.
template <typename T, int B, int MODE>
__device__ inline void op_finish(const OpArgs<T> &a, const OpCoef &cf, int64_t r, int f, uint8_t fld,
                                 typename VT<T>::t accK, typename VT<T>::t accM, typename VT<T>::t accX,
                                 const typename VT<T>::t *__restrict__ x, const typename VT<T>::t *__restrict__ b,
                                 typename VT<T>::t *__restrict__ y) {
  using V = typename VT<T>::t;
  const int64_t o = r * B + f;
  const bool fl = fld == 1;
  const V res = op_value<T>(cf, fl, accK, accM, accX);
  if (MODE == 0) {
    y[o] = res;
    return;
  }
  const V bv = b[o];
  if (MODE == 1) {
    V t; t.x = bv.x - res.x; t.y = bv.y - res.y;
    y[o] = t;
    return;
  }
  const double d = (fl ? cf.akf : cf.aks) * a.rK[r] + (fl ? cf.amf : cf.ams) * a.rM[r] + cf.ag * a.rG[r];
  const T inv = (T)(1.0 / d);
  if (MODE == 3) {
    V t; t.x = bv.x * inv; t.y = bv.y * inv;
    y[o] = t;
    return;
  }
  const V xv = x[o];
  V t;
  t.x = xv.x + (bv.x - res.x) * inv;
  t.y = xv.y + (bv.y - res.y) * inv;
  y[o] = t;
}

// Dirichlet row: the operator is the identity
template <typename T, int B, int MODE>
__device__ inline void op_identity_row(int64_t r, int f, const typename VT<T>::t *__restrict__ x,
                                       const typename VT<T>::t *__restrict__ b, typename VT<T>::t *__restrict__ y) {
  using V = typename VT<T>::t;
  const int64_t o = r * B + f;
  if (MODE == 0) y[o] = x[o];
  else if (MODE == 1) { V t; t.x = b[o].x - x[o].x; t.y = b[o].y - x[o].y; y[o] = t; }
  else y[o] = b[o];
}

// Extra entries of row r (Robin face C_G: scaled by i/(w Z); S: as is)
template <typename T, int B>
__device__ inline typename VT<T>::t op_extras(const OpArgs<T> &a, const OpCoef &cf, int64_t r, int f,
                                              const typename VT<T>::t *__restrict__ x) {
  using V = typename VT<T>::t;
  V accX = {0, 0};
  const V cg = {(T)cf.g.x, (T)cf.g.y};
  for (int32_t k = a.xrp[r]; k < a.xrp[r + 1]; k++) {
    const V xv = x[(int64_t)a.xcol[k] * B + f];
    const T v = a.xval[k];
    if (a.xtype[k] == 0) {
      V t = cmul(cg, xv);
      accX.x += v * t.x; accX.y += v * t.y;
    } else {
      accX.x += v * xv.x; accX.y += v * xv.y;
    }
  }
  return accX;
}

template <typename T, int B, int MODE>
__device__ inline void op_body(const OpArgs<T> &a, const OpCoef *__restrict__ coef,
                               const typename VT<T>::t *__restrict__ x, const typename VT<T>::t *__restrict__ b,
                               typename VT<T>::t *__restrict__ y, int blk, NarrowOut nw = {}) {
  using V = typename VT<T>::t;
  using SH = OpShape<B>;
  constexpr int G = SH::G, S = SH::S, U = SH::U, CH = SH::CH;
  __shared__ int32_t s_col[SH::RPB][CH];
  __shared__ T s_k[SH::RPB][CH];
  __shared__ T s_m[SH::RPB][CH];
  const int grp = threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const int64_t r = (int64_t)blk * SH::RPB + grp;
  if (r >= a.n) return;  // the whole group leaves together
  const int f = lane % B, sub = lane / B;
  const uint8_t fld = a.field[r];
  const OpCoef &cf = coef[f];
  if (fld == 2) {  // Dirichlet row: identity
    if (sub == 0) {
      if constexpr (MODE == 4) narrow_store(nw, r, f, x[r * B + f]);
      else op_identity_row<T, B, MODE>(r, f, x, b, y);
    }
    return;
  }
  V accK = {0, 0}, accM = {0, 0}, accX = {0, 0};
  if (MODE != 3) {
    const int32_t p0 = a.rowptr[r], p1 = a.rowptr[r + 1];
    for (int32_t base = p0; base < p1; base += CH) {
      const int cnt = min(CH, p1 - base);
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int j = u * G + lane;
        if (j < cnt) {
          s_col[grp][j] = a.colind[base + j];
          s_k[grp][j] = a.K[base + j];
          s_m[grp][j] = a.M[base + j];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 8
      for (int j = sub; j < cnt; j += S) {
        const V xv = x[(int64_t)s_col[grp][j] * B + f];
        const T kv = s_k[grp][j], mv = s_m[grp][j];
        accK.x += kv * xv.x; accK.y += kv * xv.y;
        accM.x += mv * xv.x; accM.y += mv * xv.y;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#pragma unroll
    for (int o = B; o < G; o <<= 1) {
      accK.x += __shfl_xor(accK.x, o, G); accK.y += __shfl_xor(accK.y, o, G);
      accM.x += __shfl_xor(accM.x, o, G); accM.y += __shfl_xor(accM.y, o, G);
    }
    if (sub == 0) accX = op_extras<T, B>(a, cf, r, f, x);
  }
  if (sub != 0) return;
  if constexpr (MODE == 4) narrow_store(nw, r, f, op_value<T>(cf, fld == 1, accK, accM, accX));
  else op_finish<T, B, MODE>(a, cf, r, f, fld, accK, accM, accX, x, b, y);
}

template <typename T, int B, int MODE>
__global__ __launch_bounds__(256) void op_kernel(OpArgs<T> a, const OpCoef *__restrict__ coef,
                                                 const typename VT<T>::t *__restrict__ x,
                                                 const typename VT<T>::t *__restrict__ b,
                                                 typename VT<T>::t *__restrict__ y) {
  op_body<T, B, MODE>(a, coef, x, b, y, xcd_block());
}

template <int CB>
__global__ __launch_bounds__(256) void narrow_residual_kernel(OpArgs<double> a, const OpCoef *__restrict__ coef,
                                                              const double2 *__restrict__ xn, NarrowOut nw) {
  op_body<double, CB, 4>(a, coef, xn, nullptr, nullptr, xcd_block(), nw);
}

constexpr int kRowPad = 8;  // padded rows hold a multiple of this many entries
// x of B = 32 frequencies (256 bytes per row) within 31-bit byte offsets, so
// that n * 256 fits the int and uint32 offsets of the padded operator
constexpr int64_t kRowOpMaxRows = ((1LL << 31) - 1) / 256;
static_assert(kRowOpMaxRows * 256 <= INT32_MAX, "padded operator offsets overflow");

// Row-padded copy of the FP32 operator for row_op_kernel: each row holds a
// multiple of kRowPad entries; entries store the byte offset of their column's
// x line (column * 256) and padding entries an offset past the end of x
// (which reads 0) with K = M = 0.
struct PadCsr {
  const int32_t *rp;
  const uint32_t *off;
  const float *K, *M;
};

// Row tiles of the padded FP32 operator (tile_op_kernel): consecutive rows
// (in the operator's locality order) grouped so that the distinct columns
// of a tile fit in LDS; each padded entry holds the byte offset in LDS of
// its column's line: 256 (1 + the position of the column in its tile's
// sorted column list), 0 (a zero line) for padding.
struct TileCsr {
  const int32_t *row;   // tiles + 1: first row of each tile
  const int32_t *cptr;  // tiles + 1: first column of each tile in col
  const int32_t *col;   // the tiles' distinct columns
  const uint32_t *loff; // per padded entry
  int64_t tiles;
};

// A row's l1 bound inputs (field, row sums) and a frequency's bound
// coefficients, loaded once and passed to the value functions below.
struct RowL1 {
  uint8_t fld;
  double rK, rM, rG;
};
template <typename T>
__device__ inline RowL1 row_l1(const OpArgs<T> &a, int64_t r) {
  return {a.field[r], a.rK[r], a.rM[r], a.rG[r]};
}
struct CoefL1 {
  double akf, amf, aks, ams, ag;
};
__device__ inline CoefL1 coef_l1(const OpCoef &cf) { return {cf.akf, cf.amf, cf.aks, cf.ams, cf.ag}; }

// First l1-Jacobi sweep from a zero guess, y = b / D, one thread per
// (row, frequency): the same values as op_kernel's MODE 3.
template <typename T>
__device__ inline typename VT<T>::t jacobi0_value(const RowL1 &w, const CoefL1 &cf, typename VT<T>::t bv) {
  using V = typename VT<T>::t;
  if (w.fld == 2) return bv;
  const bool fl = w.fld == 1;
  const double d = (fl ? cf.akf : cf.aks) * w.rK + (fl ? cf.amf : cf.ams) * w.rM + cf.ag * w.rG;
  const T inv = (T)(1.0 / d);
  V t;
  t.x = bv.x * inv;
  t.y = bv.y * inv;
  return t;
}
template <typename T>
__device__ inline typename VT<T>::t jacobi0_value(const OpArgs<T> &a, const OpCoef *__restrict__ coef, int64_t r,
                                                  int f, typename VT<T>::t bv) {
  return jacobi0_value<T>(row_l1<T>(a, r), coef_l1(coef[f]), bv);
}

// 1 / (l1 row bound of A(w_f)) on free rows, 0 on Dirichlet rows: the
// weights of the Jacobi-scaled residual norm ||D^-1/2 r||.  The rows of one
// field share a scale (fluid rows divided by rho w^2, structure rows of order
// E h), and fields can differ by 15 orders: a plain norm of a coupled
// residual sees the structure rows only.
__device__ inline double row_weight(const RowL1 &w, const CoefL1 &cf) {
  if (w.fld == 2) return 0.0;
  const bool fl = w.fld == 1;
  // a scale, not a value of the solution: FP32 and a hardware reciprocal
  // (about 1 ulp) are enough, and keep the update kernel memory bound
  const float d = (float)(fl ? cf.akf : cf.aks) * (float)w.rK + (float)(fl ? cf.amf : cf.ams) * (float)w.rM +
                  (float)cf.ag * (float)w.rG;
  return d > 0.f ? (double)__builtin_amdgcn_rcpf(d) : 0.0;
}
template <typename T>
__device__ inline double row_weight(const OpArgs<T> &a, const OpCoef *__restrict__ coef, int64_t r, int f) {
  return row_weight(row_l1<T>(a, r), coef_l1(coef[f]));
}

template <typename T>
__global__ __launch_bounds__(256) void jacobi0_kernel(OpArgs<T> a, int B, const OpCoef *__restrict__ coef,
                                                      const typename VT<T>::t *__restrict__ b,
                                                      typename VT<T>::t *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n * B) return;
  y[i] = jacobi0_value<T>(a, coef, i / B, (int)(i % B), b[i]);
}

// A padded row in batches of NB entries: lane u loads entry u's x offset,
// K and M for 2·NB entries with three coalesced vector loads (one memory
// round trip for both batches of a row of up to 64 entries); every entry's
// offset is broadcast with v_readlane and all NB loads of x of a batch are
// issued before its first FMA.  Entries past the row read an offset past the
// end of x (0).  Entries are summed in row order.
template <int NB>
__device__ inline void row_batches(const PadCsr &pc, int32_t base, int32_t p1, uint32_t oob,
                                   __amdgpu_buffer_rsrc_t xr, uint32_t vo, int lane, float &aK, float &aM) {
  static_assert(2 * NB == 64, "two batches per wavefront-wide index load");
  for (; base < p1; base += 2 * NB) {
    const bool in = base + lane < p1;
    const uint32_t myoff = in ? pc.off[base + lane] : oob;
    const float myk = in ? pc.K[base + lane] : 0.f, mym = in ? pc.M[base + lane] : 0.f;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (h == 1 && base + NB >= p1) break;  // uniform: the row ends in the first batch
      float xv[NB];
#pragma unroll
      for (int u = 0; u < NB; u++) {
        const uint32_t so = __builtin_amdgcn_readlane(myoff, h * NB + u);
        xv[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(xr, vo, so, 0));
      }
#pragma unroll
      for (int u = 0; u < NB; u++) {
        const int e = h * NB + u;
        aK = fmaf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, myk), e)), xv[u], aK);
        aM = fmaf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, mym), e)), xv[u], aM);
      }
    }
  }
}

// row_batches with x read from the tile's lines in LDS: the entries' LDS
// line offsets, K and M loaded 64 at a time with coalesced vector loads and
// broadcast with v_readlane
template <int NB>
__device__ inline void row_batches_lds(const PadCsr &pc, const uint32_t *__restrict__ loff, int32_t base,
                                         int32_t p1, const float *lds_lane, int lane, float &aK, float &aM) {
  static_assert(2 * NB == 64, "two batches per wavefront-wide index load");
  const char *lb = reinterpret_cast<const char *>(lds_lane);
  for (; base < p1; base += 2 * NB) {
    const bool in = base + lane < p1;
    const uint32_t myoff = in ? loff[base + lane] : 0u;
    const float myk = in ? pc.K[base + lane] : 0.f, mym = in ? pc.M[base + lane] : 0.f;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (h == 1 && base + NB >= p1) break;  // uniform: the row ends in the first batch
      float xv[NB];
#pragma unroll
      for (int u = 0; u < NB; u++)
        xv[u] = *reinterpret_cast<const float *>(lb + __builtin_amdgcn_readlane(myoff, h * NB + u));
#pragma unroll
      for (int u = 0; u < NB; u++) {
        const int e = h * NB + u;
        aK = fmaf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, myk), e)), xv[u], aK);
        aM = fmaf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, mym), e)), xv[u], aM);
      }
    }
  }
}

// Row-per-wavefront operator for B = 32 (FP32): lane l holds float l of the
// row's 32 complex frequencies (frequency l / 2, real part for even l,
// imaginary part for odd l).  K and M are real, so both parts accumulate
// independently: the row's x offsets and K, M values are the same for every
// lane (broadcast from the lane that loaded them), and each nonzero costs one
// coalesced 256-byte buffer load of x and two FMAs, with no per-entry address
// arithmetic.  The complex coefficients are applied once per row with the
// partner lane's part.  Entries are summed in row order.
constexpr int kRowWg = 256;  // threads per workgroup: kRowWg / 64 consecutive rows

template <int MODE, bool TILE = false>
__device__ inline void row_apply(const OpArgs<float> &a, const PadCsr &pc, const OpCoef *__restrict__ coef,
                                 const float *__restrict__ x, const float *__restrict__ b, float *__restrict__ y,
                                 int64_t r, int lane, const uint32_t *__restrict__ tloff = nullptr,
                                 const float *lds = nullptr) {
  constexpr int W = 64;  // floats per row: 32 complex frequencies
  const int64_t o = r * W + lane;
  const uint8_t fld = a.field[r];
  if (fld == 2) {  // Dirichlet row: identity
    if (MODE == 0) y[o] = x[o];
    else if (MODE == 1) y[o] = b[o] - x[o];
    else y[o] = b[o];
    return;
  }
  // the row's own values are loaded first so their latency overlaps the gathers
  const float bv = MODE == 0 ? 0.f : b[o];
  const float xo = MODE == 2 ? x[o] : 0.f;
  float aK = 0.f, aM = 0.f, aG = 0.f, aS = 0.f;
  if (MODE != 3) {
    const int xbytes = (int)(a.n * W * (int64_t)sizeof(float));
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, xbytes, 0x00020000);
    const uint32_t vo = (uint32_t)lane * sizeof(float);
    const int32_t base = pc.rp[r], p1 = pc.rp[r + 1];
    if constexpr (TILE) row_batches_lds<32>(pc, tloff, base, p1, lds + lane, lane, aK, aM);
    else row_batches<32>(pc, base, p1, (uint32_t)xbytes, xr, vo, lane, aK, aM);
    const float *xl = x + lane;
    for (int32_t k = a.xrp[r]; k < a.xrp[r + 1]; k++) {
      const float xv = xl[(uint32_t)a.xcol[k] * (uint32_t)W];
      if (a.xtype[k] == 0) aG = fmaf(a.xval[k], xv, aG);
      else aS = fmaf(a.xval[k], xv, aS);
    }
  }
  // partner part (real <-> imaginary) of each sum
  const float pK = __shfl_xor(aK, 1), pM = __shfl_xor(aM, 1), pG = __shfl_xor(aG, 1);
  const bool im = lane & 1;
  const float2 K = im ? float2{pK, aK} : float2{aK, pK};
  const float2 M = im ? float2{pM, aM} : float2{aM, pM};
  const float2 Gs = im ? float2{pG, aG} : float2{aG, pG};
  const OpCoef &cf = coef[lane >> 1];
  const bool fl = fld == 1;
  const float2 ck = fl ? float2{(float)cf.kf.x, (float)cf.kf.y} : float2{(float)cf.ks.x, (float)cf.ks.y};
  const float2 cm = fl ? float2{(float)cf.mf.x, (float)cf.mf.y} : float2{(float)cf.ms.x, (float)cf.ms.y};
  const float2 cg = {(float)cf.g.x, (float)cf.g.y};
  const float2 t1 = cmul(ck, K), t2 = cmul(cm, M), t3 = cmul(cg, Gs);
  const float res = im ? (t1.y + t2.y + t3.y + aS) : (t1.x + t2.x + t3.x + aS);
  if (MODE == 0) {
    y[o] = res;
    return;
  }
  if (MODE == 1) {
    y[o] = bv - res;
    return;
  }
  const double d = (fl ? cf.akf : cf.aks) * a.rK[r] + (fl ? cf.amf : cf.ams) * a.rM[r] + cf.ag * a.rG[r];
  const float inv = (float)(1.0 / d);
  if (MODE == 3) {
    y[o] = bv * inv;
    return;
  }
  y[o] = xo + (bv - res) * inv;
}

template <int MODE>
__global__ __launch_bounds__(kRowWg) void row_op_kernel(OpArgs<float> a, PadCsr pc, const OpCoef *__restrict__ coef,
                                                     const float *__restrict__ x, const float *__restrict__ b,
                                                     float *__restrict__ y) {
  const int64_t r = __builtin_amdgcn_readfirstlane(xcd_block() * (kRowWg / 64) + (int)(threadIdx.x >> 6));
  if (r >= a.n) return;
  row_apply<MODE>(a, pc, coef, x, b, y, r, threadIdx.x & 63);
}

// The row operator on row tiles: a workgroup copies the x lines of its
// tile's distinct columns into LDS once (line 0 zero), then each wavefront
// applies rows of the tile with row_apply reading x from LDS.  A row's x
// lines come from HBM/L2 once per tile instead of once per entry (a Hex20
// row has about 57 entries and neighbouring rows share most columns).
// Dynamic LDS: 256 bytes per line (32 complex FP32 frequencies).
constexpr int kTileWg = 1024;
constexpr int kTileStage = 4;  // float4 of x per thread in the staging: (255 + 1) lines x 16 / 1024
constexpr int kTileMaxLines = kTileStage * kTileWg / 16 - 1;  // 64 KB of LDS per tile with the zero line
template <int MODE>
__global__ __launch_bounds__(kTileWg) void tile_op_kernel(OpArgs<float> a, PadCsr pc, TileCsr tc,
                                                          const OpCoef *__restrict__ coef,
                                                          const float *__restrict__ x, const float *__restrict__ b,
                                                          float *__restrict__ y) {
  extern __shared__ float4 tile_lds[];
  const int64_t t = xcd_block();
  if (t >= tc.tiles) return;  // whole workgroup
  const int32_t c0 = tc.cptr[t], nl = tc.cptr[t + 1] - c0;
  const float4 *x4 = reinterpret_cast<const float4 *>(x);
  // stage: 16 float4 per line, at most kTileStage per thread; the column
  // indices first, then every x load in flight together
  const int nq = (nl + 1) * 16;
  int32_t col[kTileStage];
#pragma unroll
  for (int u = 0; u < kTileStage; u++) {
    const int i = threadIdx.x + u * kTileWg, l = i >> 4;
    col[u] = (i < nq && l > 0) ? tc.col[c0 + l - 1] : -1;
  }
  float4 v[kTileStage];
#pragma unroll
  for (int u = 0; u < kTileStage; u++) {
    const int i = threadIdx.x + u * kTileWg;
    v[u] = col[u] >= 0 ? x4[(int64_t)col[u] * 16 + (i & 15)] : float4{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int u = 0; u < kTileStage; u++)
    if (threadIdx.x + u * kTileWg < nq) tile_lds[threadIdx.x + u * kTileWg] = v[u];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const float *lds = reinterpret_cast<const float *>(tile_lds);
  const int32_t r1 = tc.row[t + 1];
  for (int32_t r = tc.row[t] + (int)(threadIdx.x >> 6); r < r1; r += kTileWg / 64)
    row_apply<MODE, true>(a, pc, coef, x, b, y, __builtin_amdgcn_readfirstlane(r), lane, tc.loff, lds);
}

// Broadcasts within each row of 16 lanes (DPP row_newbcast:J reads lane J of
// the row) folded into the instruction that uses the value.  The callers
// keep two conditions of the ISA's DPP hazards: every broadcast source (the
// first operand) is written by a memory load, never by a vector ALU
// instruction just before, and the whole wavefront is active with no EXEC
// change around these instructions.
template <int J>
__device__ inline float fmac_bcast(float acc, float k, float x) {  // acc + bcast_J(k) * x, fused
  asm("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf" : "+v"(acc) : "v"(k), "v"(x), "n"(J));
  return acc;
}
template <int J>
__device__ inline uint32_t add_bcast(uint32_t o, uint32_t v) {  // bcast_J(o) + v
  uint32_t r;
  asm("v_add_u32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf" : "=v"(r) : "v"(o), "v"(v), "n"(J));
  return r;
}
template <int J0, int N, typename Fn>
__device__ inline void static_for(Fn &&fn) {
  if constexpr (N > 0) {
    fn(std::integral_constant<int, J0>{});
    static_for<J0 + 1, N - 1>(fn);
  }
}

// x of F consecutive complex frequencies of one line: one 8- or 16-byte load
// (the whole vector is bit-cast to floats: a bit cast of one element of an
// integer vector reads element 0 with this compiler)
typedef float elp_f2v __attribute__((ext_vector_type(2)));
typedef float elp_f4v __attribute__((ext_vector_type(4)));
template <int F> struct XLine;
template <> struct XLine<1> {
  using t = elp_f2v;
  static __device__ inline t load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(t, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  }
};
template <> struct XLine<2> {
  using t = elp_f4v;
  static __device__ inline t load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(t, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  }
};

// One step's entry values of a group's row: lane j of each 16-lane row holds
// entry e + j's x offset (or LDS offset), K and M, read through range-checked
// loads (a step's second half may run past the last row: those values are
// loaded, never used).
struct GrpMeta {
  uint32_t o;
  float k, m;
};
struct GrpSrc {
  __amdgpu_buffer_rsrc_t orr, kr, mr;
  int32_t base;  // the row's first padded entry
  int j16;
};
__device__ inline GrpMeta grp_meta(const GrpSrc &g, int32_t e) {
  const uint32_t mo = (uint32_t)(g.base + e + g.j16) * 4u;
  GrpMeta t;
  t.o = __builtin_amdgcn_raw_buffer_load_b32(g.orr, mo, 0, 0);
  t.k = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(g.kr, mo, 0, 0));
  t.m = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(g.mr, mo, 0, 0));
  return t;
}

// The summing loop of a group's rows (R = 2 or 4 rows to a wavefront, F = R /
// 2 complex frequencies per lane), from the first step's entry values
// `first` (grp_meta(g, 0)): entry j's values reach the other lanes of its
// 16-lane row by row_newbcast, and LX(offset + lane bytes) loads the lane's
// 2 F floats of that entry's column.  The x loads of both halves of a step
// are in flight together (BOTH; else one half at a time, for x in LDS) and
// the next step's entry values are loaded before this step's FMAs; the
// steps alternate between two sets of those registers,
// so that no copy (a vector ALU write) ever precedes a broadcast of them.
// Entries are summed in row order.  len: the group's padded row length
// (uniform).
template <int F, bool BOTH, typename LX>
__device__ inline void grp_sum(const GrpSrc &g, const GrpMeta &first, int32_t len, uint32_t qoff, LX &&lx,
                               float (&aK)[2 * F], float (&aM)[2 * F]) {
  using XT = typename XLine<F>::t;
  auto step = [&](const GrpMeta &cur, int32_t e, GrpMeta &next) {
    XT xv[BOTH ? 16 : 8];
    const bool two = e + 8 < len;  // uniform: the step's second half is in the row
    auto ld = [&](auto J) {
      constexpr int j = decltype(J)::value;
      xv[BOTH ? j : j % 8] = lx(add_bcast<j>(cur.o, qoff));
    };
    auto fm = [&](auto J) {
      constexpr int j = decltype(J)::value;
#pragma unroll
      for (int c = 0; c < 2 * F; c++) {
        aK[c] = fmac_bcast<j>(aK[c], cur.k, xv[BOTH ? j : j % 8][c]);
        aM[c] = fmac_bcast<j>(aM[c], cur.m, xv[BOTH ? j : j % 8][c]);
      }
    };
    static_for<0, 8>(ld);
    if (BOTH && two) static_for<8, 8>(ld);
    if (e + 16 < len) next = grp_meta(g, e + 16);
    static_for<0, 8>(fm);
    if (two) {
      if (!BOTH) static_for<8, 8>(ld);
      static_for<8, 8>(fm);
    }
  };
  if (len > 0) {
    GrpMeta ma = first, mb;
    for (int32_t e = 0;;) {
      step(ma, e, mb);
      if ((e += 16) >= len) break;
      step(mb, e, ma);
      if ((e += 16) >= len) break;
    }
  }
}

// A group row's own values for the epilogue, loaded before its sum so that
// their latency overlaps it: field, extra-entry range, b (MODE 1, 2), x
// (MODE 2; MODE 0 and 1 read x in the epilogue for Dirichlet rows only).
template <int MODE, int F>
struct GrpRow {
  uint8_t fld;
  int32_t x0, x1;
  float bv[2 * F], xo[2 * F];
};
template <int MODE, int F>
__device__ inline GrpRow<MODE, F> grp_row(const OpArgs<float> &a, const float *__restrict__ x,
                                          const float *__restrict__ b, int32_t r, int q) {
  constexpr int W = 64;
  GrpRow<MODE, F> w;
  const int64_t o = (int64_t)r * W + q * 2 * F;
  w.fld = a.field[r];
  w.x0 = a.xrp[r];
  w.x1 = a.xrp[r + 1];
#pragma unroll
  for (int c = 0; c < 2 * F; c++) {
    w.bv[c] = MODE == 0 ? 0.f : b[o + c];
    w.xo[c] = MODE == 2 ? x[o + c] : 0.f;
  }
  return w;
}

// Row epilogue of the grouped kernels (row_apply's, for the lane's F
// frequencies q F .. q F + F - 1): extra entries (Robin face, FSI), the
// coefficients, MODE, Dirichlet rows as identity.
template <int MODE, int F>
__device__ inline void grp_finish(const OpArgs<float> &a, const OpCoef *__restrict__ coef,
                                  const float *__restrict__ x, float *__restrict__ y, int32_t r, int q,
                                  const GrpRow<MODE, F> &w, const float (&aK)[2 * F], const float (&aM)[2 * F]) {
  constexpr int W = 64;
  const uint8_t fld = w.fld;
  const int64_t o = (int64_t)r * W + q * 2 * F;
  float aG[2 * F], aS[2 * F], xo[2 * F];
  const float *bv = w.bv;
#pragma unroll
  for (int c = 0; c < 2 * F; c++) {
    aG[c] = aS[c] = 0.f;
    xo[c] = MODE == 2 ? w.xo[c] : ((MODE == 0 || MODE == 1) && fld == 2) ? x[o + c] : 0.f;
  }
  for (int32_t k = w.x0; k < w.x1; k++) {
    const float *xl = x + (int64_t)(uint32_t)a.xcol[k] * W + q * 2 * F;
    const float v = a.xval[k];
    const bool g = a.xtype[k] == 0;
#pragma unroll
    for (int c = 0; c < 2 * F; c++) {
      if (g) aG[c] = fmaf(v, xl[c], aG[c]);
      else aS[c] = fmaf(v, xl[c], aS[c]);
    }
  }
  const bool fl = fld == 1;
  float res[2 * F];
#pragma unroll
  for (int t = 0; t < F; t++) {
    const OpCoef &cf = coef[q * F + t];
    const float2 K = {aK[2 * t], aK[2 * t + 1]}, M = {aM[2 * t], aM[2 * t + 1]};
    const float2 Gs = {aG[2 * t], aG[2 * t + 1]};
    const float2 ck = fl ? float2{(float)cf.kf.x, (float)cf.kf.y} : float2{(float)cf.ks.x, (float)cf.ks.y};
    const float2 cm = fl ? float2{(float)cf.mf.x, (float)cf.mf.y} : float2{(float)cf.ms.x, (float)cf.ms.y};
    const float2 cg = {(float)cf.g.x, (float)cf.g.y};
    const float2 t1 = cmul(ck, K), t2 = cmul(cm, M), t3 = cmul(cg, Gs);
    const float re = t1.x + t2.x + t3.x + aS[2 * t], im = t1.y + t2.y + t3.y + aS[2 * t + 1];
    if (MODE == 0) {
      res[2 * t] = re;
      res[2 * t + 1] = im;
    } else if (MODE == 1) {
      res[2 * t] = bv[2 * t] - re;
      res[2 * t + 1] = bv[2 * t + 1] - im;
    } else {
      const double d = (fl ? cf.akf : cf.aks) * a.rK[r] + (fl ? cf.amf : cf.ams) * a.rM[r] + cf.ag * a.rG[r];
      const float inv = (float)(1.0 / d);
      if (MODE == 3) {
        res[2 * t] = bv[2 * t] * inv;
        res[2 * t + 1] = bv[2 * t + 1] * inv;
      } else {
        res[2 * t] = xo[2 * t] + (bv[2 * t] - re) * inv;
        res[2 * t + 1] = xo[2 * t + 1] + (bv[2 * t + 1] - im) * inv;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 2 * F; c++) {
    float v = res[c];
    if (fld == 2) v = MODE == 0 ? xo[c] : MODE == 1 ? bv[c] - xo[c] : bv[c];  // Dirichlet row: identity
    y[o + c] = v;
  }
}

// The 32-slot FP32 operator with R rows to a wavefront (R = 2 or 4): each row
// is taken by G = 64 / R lanes, lane q of a row holding F = R / 2 consecutive
// complex frequencies (2 F floats, one 8- or 16-byte load of x per entry).
// The rows of a wavefront (a group: rows of one window of 64 consecutive
// rows taken by length) are padded to one length, so the wavefront walks
// their entries in lock step with every lane active (grp_sum).  Per entry
// and row: one load of x, 4 F FMAs and one address add for the wavefront's
// R rows, against three v_readlane broadcasts and two FMAs per entry in
// row_op_kernel.  Entries are summed in row order and the row epilogue is
// row_apply's, so the results are the same bits.  `grows` lists each group's
// R rows (-1 past the last row; a group's first row is always present).
template <int MODE, int R>
__global__ __launch_bounds__(kRowWg) void rgrp_op_kernel(OpArgs<float> a, PadCsr pc, int64_t n_pad,
                                                       const int32_t *__restrict__ grows, int64_t waves,
                                                       const OpCoef *__restrict__ coef, const float *__restrict__ x,
                                                       const float *__restrict__ b, float *__restrict__ y) {
  static_assert(R == 2 || R == 4, "2 or 4 rows per wavefront");
  constexpr int G = 64 / R, F = R / 2, W = 64;
  const int64_t wv = __builtin_amdgcn_readfirstlane(xcd_block() * (kRowWg / 64) + (int)(threadIdx.x >> 6));
  if (wv >= waves) return;
  const int lane = threadIdx.x & 63, q = lane % G;
  const int32_t r0 = grows[wv * R + lane / G];
  const int32_t rf = __builtin_amdgcn_readfirstlane(r0);
  const int32_t r = r0 >= 0 ? r0 : rf;  // an absent row walks the first row's entries, unstored
  const int32_t len = MODE == 3 ? 0 : __builtin_amdgcn_readfirstlane(pc.rp[rf + 1] - pc.rp[rf]);
  const int xbytes = (int)(a.n * W * (int64_t)sizeof(float));
  const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, xbytes, 0x00020000);
  const int ebytes = (int)(n_pad * (int64_t)sizeof(float));
  const __amdgpu_buffer_rsrc_t orr = __builtin_amdgcn_make_buffer_rsrc((void *)pc.off, (short)0, ebytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t kr = __builtin_amdgcn_make_buffer_rsrc((void *)pc.K, (short)0, ebytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t mr = __builtin_amdgcn_make_buffer_rsrc((void *)pc.M, (short)0, ebytes, 0x00020000);
  const GrpSrc gs{orr, kr, mr, pc.rp[r], lane & 15};
  const GrpMeta first = grp_meta(gs, 0);
  const GrpRow<MODE, F> w = grp_row<MODE, F>(a, x, b, r, q);
  float aK[2 * F], aM[2 * F];
#pragma unroll
  for (int c = 0; c < 2 * F; c++) aK[c] = aM[c] = 0.f;
  grp_sum<F, true>(gs, first, len, (uint32_t)q * (8 * F), [&](uint32_t off) { return XLine<F>::load(xr, off); },
                   aK, aM);
  if (r0 < 0) return;
  grp_finish<MODE, F>(a, coef, x, y, r, q, w, aK, aM);
}

// The grouped operator on row tiles (tgrp_op_kernel): a workgroup of kTgWg
// threads stages the x lines of its tile's distinct columns in LDS (line 0
// zero), then each wavefront takes groups of R rows of the tile (rows of the
// tile taken by length, padded to one length per group) and sums them with
// grp_sum reading x from LDS: per padded entry the tile's LDS byte offset of
// its column's line (256 (1 + position), 0 for padding) replaces the offset
// in x.  Each x line crosses the texture path once per tile instead of once
// per entry.  Same entries, same order: the same bits.
constexpr int kTgWg = 512;
template <int MODE, int R>
__global__ __launch_bounds__(kTgWg) void tgrp_op_kernel(OpArgs<float> a, PadCsr pc, int64_t n_pad, TileCsr tc,
                                                      const int32_t *__restrict__ tgrp,
                                                      const int32_t *__restrict__ grows,
                                                      const OpCoef *__restrict__ coef, const float *__restrict__ x,
                                                      const float *__restrict__ b, float *__restrict__ y) {
  static_assert(R == 2 || R == 4, "2 or 4 rows per wavefront");
  constexpr int G = 64 / R, F = R / 2, ST = 8;  // ST float4 of x per thread per staging round
  extern __shared__ float4 tile_lds[];
  const int64_t t = xcd_block();
  if (t >= tc.tiles) return;  // whole workgroup
  const int32_t c0 = tc.cptr[t], nl = tc.cptr[t + 1] - c0;
  const float4 *x4 = reinterpret_cast<const float4 *>(x);
  const int nq = (nl + 1) * 16;
  const int lane = threadIdx.x & 63, q = lane % G;
  const int ebytes = (int)(n_pad * (int64_t)sizeof(float));
  const __amdgpu_buffer_rsrc_t orr = __builtin_amdgcn_make_buffer_rsrc((void *)tc.loff, (short)0, ebytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t kr = __builtin_amdgcn_make_buffer_rsrc((void *)pc.K, (short)0, ebytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t mr = __builtin_amdgcn_make_buffer_rsrc((void *)pc.M, (short)0, ebytes, 0x00020000);
  // the wavefront's first group: its rows, entry values and row values are
  // loaded before the staging, so that their latency overlaps it
  const int32_t g1 = tgrp[t + 1];
  int32_t gi = tgrp[t] + (int)(threadIdx.x >> 6);
  int32_t r0 = -1, r = 0, len = 0;
  GrpSrc gs{orr, kr, mr, 0, lane & 15};
  GrpMeta first{};
  GrpRow<MODE, F> w{};
  auto open_group = [&]() {
    const int32_t g = __builtin_amdgcn_readfirstlane(gi);
    r0 = grows[(int64_t)g * R + lane / G];
    const int32_t rf = __builtin_amdgcn_readfirstlane(r0);
    r = r0 >= 0 ? r0 : rf;
    len = MODE == 3 ? 0 : __builtin_amdgcn_readfirstlane(pc.rp[rf + 1] - pc.rp[rf]);
    gs.base = pc.rp[r];
    first = grp_meta(gs, 0);
    w = grp_row<MODE, F>(a, x, b, r, q);
  };
  if (gi < g1) open_group();
  for (int i0 = 0; i0 < nq; i0 += ST * kTgWg) {  // one round for up to 255 lines
    int32_t col[ST];
#pragma unroll
    for (int u = 0; u < ST; u++) {
      const int i = i0 + threadIdx.x + u * kTgWg, l = i >> 4;
      col[u] = (i < nq && l > 0) ? tc.col[c0 + l - 1] : -1;
    }
    float4 v[ST];
#pragma unroll
    for (int u = 0; u < ST; u++) {
      const int i = i0 + threadIdx.x + u * kTgWg;
      v[u] = col[u] >= 0 ? x4[(int64_t)col[u] * 16 + (i & 15)] : float4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < ST; u++) {
      const int i = i0 + threadIdx.x + u * kTgWg;
      if (i < nq) tile_lds[i] = v[u];
    }
  }
  __syncthreads();
  const char *lb = reinterpret_cast<const char *>(tile_lds);
  while (gi < g1) {
    float aK[2 * F], aM[2 * F];
#pragma unroll
    for (int c = 0; c < 2 * F; c++) aK[c] = aM[c] = 0.f;
    grp_sum<F, false>(gs, first, len, (uint32_t)q * (8 * F),
                      [&](uint32_t off) { return *reinterpret_cast<const typename XLine<F>::t *>(lb + off); }, aK,
                      aM);
    if (r0 >= 0) grp_finish<MODE, F>(a, coef, x, y, r, q, w, aK, aM);
    if ((gi += kTgWg / 64) < g1) open_group();
  }
}

// tgrp_op_kernel with each tile's 32 frequencies split over two workgroups
// (thgrp_op_kernel): workgroup 2 t + h stages the half-lines (128 bytes,
// frequencies 16 h .. 16 h + 15) of tile t's columns, half the LDS of a
// whole-line tile, so twice as many workgroups fit a CU; each 16-lane row of
// a wavefront holds one of its row's 16 frequencies per lane (8-byte LDS
// reads), and the tile's entries are read by both workgroups (the second
// read from L2); the entries' LDS offsets count 128-byte lines.  Same
// entries, same order: the same bits.
template <int MODE>
__global__ __launch_bounds__(kTgWg) void thgrp_op_kernel(OpArgs<float> a, PadCsr pc, int64_t n_pad, TileCsr tc,
                                                       const int32_t *__restrict__ tgrp,
                                                       const int32_t *__restrict__ grows,
                                                       const OpCoef *__restrict__ coef, const float *__restrict__ x,
                                                       const float *__restrict__ b, float *__restrict__ y) {
  constexpr int R = 4, G = 16, F = 1, ST = 4;  // ST float4 of x per thread per staging round
  extern __shared__ float4 tile_lds[];
  const int64_t tb = xcd_block();
  const int64_t t = tb >> 1;
  const int h = (int)(tb & 1);
  if (t >= tc.tiles) return;  // whole workgroup
  const int32_t c0 = tc.cptr[t], nl = tc.cptr[t + 1] - c0;
  const float4 *x4 = reinterpret_cast<const float4 *>(x);
  const int nq = (nl + 1) * 8;
  const int lane = threadIdx.x & 63, q = 16 * h + lane % G;  // the lane's frequency
  const int ebytes = (int)(n_pad * (int64_t)sizeof(float));
  const __amdgpu_buffer_rsrc_t orr = __builtin_amdgcn_make_buffer_rsrc((void *)tc.loff, (short)0, ebytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t kr = __builtin_amdgcn_make_buffer_rsrc((void *)pc.K, (short)0, ebytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t mr = __builtin_amdgcn_make_buffer_rsrc((void *)pc.M, (short)0, ebytes, 0x00020000);
  const int32_t g1 = tgrp[t + 1];
  int32_t gi = tgrp[t] + (int)(threadIdx.x >> 6);
  int32_t r0 = -1, r = 0, len = 0;
  GrpSrc gs{orr, kr, mr, 0, lane & 15};
  GrpMeta first{};
  GrpRow<MODE, F> w{};
  auto open_group = [&]() {
    const int32_t g = __builtin_amdgcn_readfirstlane(gi);
    r0 = grows[(int64_t)g * R + lane / G];
    const int32_t rf = __builtin_amdgcn_readfirstlane(r0);
    r = r0 >= 0 ? r0 : rf;
    len = MODE == 3 ? 0 : __builtin_amdgcn_readfirstlane(pc.rp[rf + 1] - pc.rp[rf]);
    gs.base = pc.rp[r];
    first = grp_meta(gs, 0);
    w = grp_row<MODE, F>(a, x, b, r, q);
  };
  if (gi < g1) open_group();
  for (int i0 = 0; i0 < nq; i0 += ST * kTgWg) {  // one round for up to 255 lines
    int32_t col[ST];
#pragma unroll
    for (int u = 0; u < ST; u++) {
      const int i = i0 + threadIdx.x + u * kTgWg, l = i >> 3;
      col[u] = (i < nq && l > 0) ? tc.col[c0 + l - 1] : -1;
    }
    float4 v[ST];
#pragma unroll
    for (int u = 0; u < ST; u++) {
      const int i = i0 + threadIdx.x + u * kTgWg;
      v[u] = col[u] >= 0 ? x4[(int64_t)col[u] * 16 + 8 * h + (i & 7)] : float4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < ST; u++) {
      const int i = i0 + threadIdx.x + u * kTgWg;
      if (i < nq) tile_lds[i] = v[u];
    }
  }
  __syncthreads();
  const char *lb = reinterpret_cast<const char *>(tile_lds);
  while (gi < g1) {
    float aK[2 * F], aM[2 * F];
#pragma unroll
    for (int c = 0; c < 2 * F; c++) aK[c] = aM[c] = 0.f;
    grp_sum<F, false>(gs, first, len, (uint32_t)(lane % G) * 8u,
                      [&](uint32_t off) { return *reinterpret_cast<const typename XLine<F>::t *>(lb + off); }, aK,
                      aM);
    if (r0 >= 0) grp_finish<MODE, F>(a, coef, x, y, r, q, w, aK, aM);
    if ((gi += kTgWg / 64) < g1) open_group();
  }
}

// Post-smoothing of the two-level box cycle fused with the prolongation,
// for 32 frequencies.  With u = D^-1 r (first sweep), v = r - A u and e the
// box-lattice correction, the cycle's last step is
//   z = x + D^-1 (r - A x),  x = u + P e,
// and A x = (r - v) + (A P) e, so
//   z = u + P e + D^-1 (v - (cK KP + cM MP + cG GP) e)
// with the frequency-independent products KP, MP, GP (C_Gamma P) formed once.
// A Hex20 row of A P has at most 27 lattice columns (57 for A), and e is the
// lattice vector, a quarter of the fine one.  Row layout as row_op_kernel:
// one wavefront per row, lane = frequency x real/imaginary part, rows padded
// to multiples of kRowPad with offsets past the end of e; P rows padded to 8.
struct PadCsr3 {
  const int32_t *rp;
  const uint32_t *off;
  const float *K, *M, *G;
};

// (A P) e over a padded row in batches of NB entries (row_batches' scheme:
// per-entry values loaded by the lanes, broadcast with v_readlane)
template <int NB>
__device__ inline void ap_batches(const PadCsr3 &pc, int32_t base, int32_t p1, uint32_t oob,
                                  __amdgpu_buffer_rsrc_t er, uint32_t vo, int lane, float &aK, float &aM, float &aG) {
  for (; base < p1; base += NB) {
    const bool in = lane < NB && base + lane < p1;
    const uint32_t myoff = in ? pc.off[base + lane] : oob;
    const float myk = in ? pc.K[base + lane] : 0.f, mym = in ? pc.M[base + lane] : 0.f;
    const float myg = in ? pc.G[base + lane] : 0.f;
    float xv[NB];
#pragma unroll
    for (int u = 0; u < NB; u++) {
      const uint32_t so = __builtin_amdgcn_readlane(myoff, u);
      xv[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(er, vo, so, 0));
    }
#pragma unroll
    for (int u = 0; u < NB; u++) {
      aK = fmaf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, myk), u)), xv[u], aK);
      aM = fmaf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, mym), u)), xv[u], aM);
      aG = fmaf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, myg), u)), xv[u], aG);
    }
  }
}

__global__ __launch_bounds__(kRowWg) void box_post_kernel(OpArgs<float> a, PadCsr3 ap, const uint32_t *__restrict__ poff,
                                                       const float *__restrict__ pval, const OpCoef *__restrict__ coef,
                                                       const float *__restrict__ e, int64_t ne,
                                                       const float *__restrict__ u, const float *__restrict__ v,
                                                       float *__restrict__ z) {
  constexpr int W = 64;
  const int lane = threadIdx.x & 63;
  const int64_t r = __builtin_amdgcn_readfirstlane(xcd_block() * (kRowWg / 64) + (int)(threadIdx.x >> 6));
  if (r >= a.n) return;
  const int64_t o = r * W + lane;
  const float uu = u[o], vv = v[o];
  const int ebytes = (int)(ne * W * (int64_t)sizeof(float));
  const __amdgpu_buffer_rsrc_t er = __builtin_amdgcn_make_buffer_rsrc((void *)e, (short)0, ebytes, 0x00020000);
  const uint32_t vo = (uint32_t)lane * sizeof(float);
  // P e
  float pe = 0.f;
  {
    uint32_t so[8];
    float pv[8], xv[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
      so[t] = poff[r * 8 + t];
      pv[t] = pval[r * 8 + t];
    }
#pragma unroll
    for (int t = 0; t < 8; t++) xv[t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(er, vo, so[t], 0));
#pragma unroll
    for (int t = 0; t < 8; t++) pe = fmaf(pv[t], xv[t], pe);
  }
  // (A P) e
  float aK = 0.f, aM = 0.f, aG = 0.f;
  ap_batches<32>(ap, ap.rp[r], ap.rp[r + 1], (uint32_t)ebytes, er, vo, lane, aK, aM, aG);
  const float pK = __shfl_xor(aK, 1), pM = __shfl_xor(aM, 1), pG = __shfl_xor(aG, 1);
  const bool im = lane & 1;
  const float2 K = im ? float2{pK, aK} : float2{aK, pK};
  const float2 M = im ? float2{pM, aM} : float2{aM, pM};
  const float2 G = im ? float2{pG, aG} : float2{aG, pG};
  const OpCoef &cf = coef[lane >> 1];
  const bool fl = a.field[r] == 1;
  const float2 ck = fl ? float2{(float)cf.kf.x, (float)cf.kf.y} : float2{(float)cf.ks.x, (float)cf.ks.y};
  const float2 cm = fl ? float2{(float)cf.mf.x, (float)cf.mf.y} : float2{(float)cf.ms.x, (float)cf.ms.y};
  const float2 cg = {(float)cf.g.x, (float)cf.g.y};
  const float2 t1 = cmul(ck, K), t2 = cmul(cm, M), t3 = cmul(cg, G);
  const float ax = im ? (t1.y + t2.y + t3.y) : (t1.x + t2.x + t3.x);
  const double d = (fl ? cf.akf : cf.aks) * a.rK[r] + (fl ? cf.amf : cf.ams) * a.rM[r] + cf.ag * a.rG[r];
  const float inv = (float)(1.0 / d);
  z[o] = uu + pe + (vv - ax) * inv;
}

// box_post_kernel on grouped half-line tiles (thbox_post_kernel), the
// scheme of thgrp_op_kernel: a row's entries are its (A P) row with the P
// value of each column beside K P, M P and C_Gamma P (0 where P has no
// entry: every column of a P row is a column of its (A P) row), so P e and
// (A P) e are summed in one walk over the lattice lines of the tile staged in
// LDS, in the column order of both rows: the same bits as box_post_kernel.
struct BoxTiles {
  const int32_t *rp;     // padded rows (per group of 4 of a tile)
  const uint32_t *loff;  // LDS offset of the entry's 128-byte half-line
  const float *K, *M, *G, *P;
  int64_t n_pad;
  TileCsr tc;            // tiles: rows, distinct lattice columns
  const int32_t *tgrp, *grows;
};
struct GrpMeta4 {
  uint32_t o;
  float k, m, g, p;
};
struct GrpSrc4 {
  __amdgpu_buffer_rsrc_t orr, kr, mr, gr, pr;
  int32_t base;
  int j16;
};
__device__ inline GrpMeta4 grp_meta4(const GrpSrc4 &g, int32_t e) {
  const uint32_t mo = (uint32_t)(g.base + e + g.j16) * 4u;
  GrpMeta4 t;
  t.o = __builtin_amdgcn_raw_buffer_load_b32(g.orr, mo, 0, 0);
  t.k = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(g.kr, mo, 0, 0));
  t.m = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(g.mr, mo, 0, 0));
  t.g = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(g.gr, mo, 0, 0));
  t.p = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(g.pr, mo, 0, 0));
  return t;
}
// grp_sum with four values per entry (K P, M P, C_Gamma P, P), one complex
// frequency per lane, one half of a step's x loads in flight at a time
template <typename LX>
__device__ inline void grp_sum4(const GrpSrc4 &g, const GrpMeta4 &first, int32_t len, uint32_t qoff, LX &&lx,
                                float (&aK)[2], float (&aM)[2], float (&aG)[2], float (&aP)[2]) {
  using XT = typename XLine<1>::t;
  auto step = [&](const GrpMeta4 &cur, int32_t e, GrpMeta4 &next) {
    XT xv[8];
    const bool two = e + 8 < len;  // uniform
    auto ld = [&](auto J) {
      constexpr int j = decltype(J)::value;
      xv[j % 8] = lx(add_bcast<j>(cur.o, qoff));
    };
    auto fm = [&](auto J) {
      constexpr int j = decltype(J)::value;
#pragma unroll
      for (int c = 0; c < 2; c++) {
        aK[c] = fmac_bcast<j>(aK[c], cur.k, xv[j % 8][c]);
        aM[c] = fmac_bcast<j>(aM[c], cur.m, xv[j % 8][c]);
        aG[c] = fmac_bcast<j>(aG[c], cur.g, xv[j % 8][c]);
        aP[c] = fmac_bcast<j>(aP[c], cur.p, xv[j % 8][c]);
      }
    };
    static_for<0, 8>(ld);
    if (e + 16 < len) next = grp_meta4(g, e + 16);
    static_for<0, 8>(fm);
    if (two) {
      static_for<8, 8>(ld);
      static_for<8, 8>(fm);
    }
  };
  if (len > 0) {
    GrpMeta4 ma = first, mb;
    for (int32_t e = 0;;) {
      step(ma, e, mb);
      if ((e += 16) >= len) break;
      step(mb, e, ma);
      if ((e += 16) >= len) break;
    }
  }
}

__global__ __launch_bounds__(kTgWg) void thbox_post_kernel(OpArgs<float> a, BoxTiles bt,
                                                         const OpCoef *__restrict__ coef,
                                                         const float *__restrict__ e, const float *__restrict__ u,
                                                         const float *__restrict__ v, float *__restrict__ z) {
  constexpr int R = 4, G = 16, ST = 4, W = 64;
  extern __shared__ float4 tile_lds[];
  const TileCsr &tc = bt.tc;
  const int64_t tb = xcd_block();
  const int64_t t = tb >> 1;
  const int h = (int)(tb & 1);
  if (t >= tc.tiles) return;  // whole workgroup
  const int32_t c0 = tc.cptr[t], nl = tc.cptr[t + 1] - c0;
  const float4 *e4 = reinterpret_cast<const float4 *>(e);
  const int nq = (nl + 1) * 8;
  const int lane = threadIdx.x & 63, q = 16 * h + lane % G;  // the lane's frequency
  const int ebytes = (int)(bt.n_pad * (int64_t)sizeof(float));
  GrpSrc4 gs;
  gs.orr = __builtin_amdgcn_make_buffer_rsrc((void *)bt.loff, (short)0, ebytes, 0x00020000);
  gs.kr = __builtin_amdgcn_make_buffer_rsrc((void *)bt.K, (short)0, ebytes, 0x00020000);
  gs.mr = __builtin_amdgcn_make_buffer_rsrc((void *)bt.M, (short)0, ebytes, 0x00020000);
  gs.gr = __builtin_amdgcn_make_buffer_rsrc((void *)bt.G, (short)0, ebytes, 0x00020000);
  gs.pr = __builtin_amdgcn_make_buffer_rsrc((void *)bt.P, (short)0, ebytes, 0x00020000);
  gs.base = 0;
  gs.j16 = lane & 15;
  const int32_t g1 = bt.tgrp[t + 1];
  int32_t gi = bt.tgrp[t] + (int)(threadIdx.x >> 6);
  int32_t r0 = -1, r = 0, len = 0;
  GrpMeta4 first{};
  float2 uu{}, vv{};
  auto open_group = [&]() {
    const int32_t g = __builtin_amdgcn_readfirstlane(gi);
    r0 = bt.grows[(int64_t)g * R + lane / G];
    const int32_t rf = __builtin_amdgcn_readfirstlane(r0);
    r = r0 >= 0 ? r0 : rf;
    len = __builtin_amdgcn_readfirstlane(bt.rp[rf + 1] - bt.rp[rf]);
    gs.base = bt.rp[r];
    first = grp_meta4(gs, 0);
    const int64_t o = (int64_t)r * W + 2 * q;
    uu = float2{u[o], u[o + 1]};
    vv = float2{v[o], v[o + 1]};
  };
  if (gi < g1) open_group();
  for (int i0 = 0; i0 < nq; i0 += ST * kTgWg) {  // one round for up to 255 lines
    int32_t col[ST];
#pragma unroll
    for (int k = 0; k < ST; k++) {
      const int i = i0 + threadIdx.x + k * kTgWg, l = i >> 3;
      col[k] = (i < nq && l > 0) ? tc.col[c0 + l - 1] : -1;
    }
    float4 x[ST];
#pragma unroll
    for (int k = 0; k < ST; k++) {
      const int i = i0 + threadIdx.x + k * kTgWg;
      x[k] = col[k] >= 0 ? e4[(int64_t)col[k] * 16 + 8 * h + (i & 7)] : float4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int k = 0; k < ST; k++) {
      const int i = i0 + threadIdx.x + k * kTgWg;
      if (i < nq) tile_lds[i] = x[k];
    }
  }
  __syncthreads();
  const char *lb = reinterpret_cast<const char *>(tile_lds);
  while (gi < g1) {
    float aK[2] = {0.f, 0.f}, aM[2] = {0.f, 0.f}, aG[2] = {0.f, 0.f}, aP[2] = {0.f, 0.f};
    grp_sum4(gs, first, len, (uint32_t)(lane % G) * 8u,
             [&](uint32_t off) { return *reinterpret_cast<const typename XLine<1>::t *>(lb + off); }, aK, aM, aG,
             aP);
    if (r0 >= 0) {
      // box_post_kernel's epilogue for the lane's frequency q
      const float2 K = {aK[0], aK[1]}, M = {aM[0], aM[1]}, Gs = {aG[0], aG[1]};
      const OpCoef &cf = coef[q];
      const bool fl = a.field[r] == 1;
      const float2 ck = fl ? float2{(float)cf.kf.x, (float)cf.kf.y} : float2{(float)cf.ks.x, (float)cf.ks.y};
      const float2 cm = fl ? float2{(float)cf.mf.x, (float)cf.mf.y} : float2{(float)cf.ms.x, (float)cf.ms.y};
      const float2 cg = {(float)cf.g.x, (float)cf.g.y};
      const float2 t1 = cmul(ck, K), t2 = cmul(cm, M), t3 = cmul(cg, Gs);
      const float axr = t1.x + t2.x + t3.x, axi = t1.y + t2.y + t3.y;
      const double d = (fl ? cf.akf : cf.aks) * a.rK[r] + (fl ? cf.amf : cf.ams) * a.rM[r] + cf.ag * a.rG[r];
      const float inv = (float)(1.0 / d);
      const int64_t o = (int64_t)r * W + 2 * q;
      z[o] = uu.x + aP[0] + (vv.x - axr) * inv;
      z[o + 1] = uu.y + aP[1] + (vv.y - axi) * inv;
    }
    if ((gi += kTgWg / 64) < g1) open_group();
  }
}

// y = P x (ADD = false) or y += P x (ADD = true) for a real CSR P (same
// chunked row walk as the operator)
template <typename T, int B, bool ADD>
__global__ __launch_bounds__(256) void transfer_kernel(int64_t n, const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ ci, const T *__restrict__ v,
                                                       const typename VT<T>::t *__restrict__ x,
                                                       typename VT<T>::t *__restrict__ y) {
  using V = typename VT<T>::t;
  using SH = OpShape<B>;
  constexpr int G = SH::G, S = SH::S, U = SH::U, CH = SH::CH;
  __shared__ int32_t s_col[SH::RPB][CH];
  __shared__ T s_v[SH::RPB][CH];
  const int grp = threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const int64_t r = (int64_t)xcd_block() * SH::RPB + grp;
  if (r >= n) return;
  const int f = lane % B, sub = lane / B;
  V acc = {0, 0};
  const int32_t p0 = rp[r], p1 = rp[r + 1];
  for (int32_t base = p0; base < p1; base += CH) {
    const int cnt = min(CH, p1 - base);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = u * G + lane;
      if (j < cnt) {
        s_col[grp][j] = ci[base + j];
        s_v[grp][j] = v[base + j];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 8
    for (int j = sub; j < cnt; j += S) {
      const V xv = x[(int64_t)s_col[grp][j] * B + f];
      const T w = s_v[grp][j];
      acc.x += w * xv.x;
      acc.y += w * xv.y;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
#pragma unroll
  for (int o = B; o < G; o <<= 1) {
    acc.x += __shfl_xor(acc.x, o, G);
    acc.y += __shfl_xor(acc.y, o, G);
  }
  if (sub != 0) return;
  const int64_t o = r * B + f;
  if (ADD) {
    V t = y[o];
    t.x += acc.x;
    t.y += acc.y;
    y[o] = t;
  } else {
    y[o] = acc;
  }
}

// Deterministic per-frequency reductions over [n][B] vectors.
//   mode 0: sum a*b (unconjugated), mode 1: sum |a|^2
// Block k of kRedBlocks sums rows [k*chunk, (k+1)*chunk) with lane (f, rl)
// taking rows rl, rl + R, ... (R = 256 / B) in order, four loads in flight;
// the per-block partials are combined by dot_final_kernel in a fixed tree.
constexpr int kRedBlocks = 2048;
// blocks of a reduction over n rows: at most kRedBlocks, at least 64 rows
// per block (small models would otherwise spend the reduction on a final
// sum over thousands of near-empty partials); a fixed function of n, so the
// partition, and with it every sum, is the same for the same model
static inline int red_blocks(int64_t n) {
  return (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(8, (n + 63) / 64));
}
constexpr int kRedFinalThreads = 1024;

template <typename V>
__device__ inline void red_term(double2 &acc, V av, V bv, int mode) {
  if (mode == 0) {
    acc.x += (double)av.x * bv.x - (double)av.y * bv.y;
    acc.y += (double)av.x * bv.y + (double)av.y * bv.x;
  } else {
    acc.x += (double)av.x * av.x + (double)av.y * av.y;
  }
}

// block tree over the R row-lanes of each frequency; lane rl == 0 writes
__device__ inline void red_block_store(double2 acc, int B, double2 *__restrict__ partial, int blk) {
  __shared__ double2 sh[256];
  const int t = threadIdx.x, rl = t / B, R = 256 / B;
  sh[t] = acc;
  __syncthreads();
  for (int st = R / 2; st > 0; st >>= 1) {
    if (rl < st) {
      sh[t].x += sh[t + st * B].x;
      sh[t].y += sh[t + st * B].y;
    }
    __syncthreads();
  }
  if (rl == 0) partial[blk * B + t] = sh[t];
}

template <typename V>
__device__ inline void dot_body(int64_t n, int B, const V *__restrict__ a, const V *__restrict__ b, int mode,
                                double2 *__restrict__ partial, int blk, int nblocks) {
  const int t = threadIdx.x, f = t % B, rl = t / B, R = 256 / B;
  const int64_t chunk = (n + nblocks - 1) / nblocks;
  const int64_t r0 = blk * chunk, r1 = min(n, r0 + chunk);
  double2 acc = {0, 0};
  int64_t r = r0 + rl;
  for (; r + 3 * R < r1; r += 4 * R) {
    V av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      av[u] = a[(r + u * R) * B + f];
      bv[u] = mode == 0 ? b[(r + u * R) * B + f] : av[u];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) red_term(acc, av[u], bv[u], mode);
  }
  for (; r < r1; r += R) red_term(acc, a[r * B + f], mode == 0 ? b[r * B + f] : a[r * B + f], mode);
  red_block_store(acc, B, partial, blk);
}

template <typename V>
__global__ __launch_bounds__(256) void dot_kernel(int64_t n, int B, const V *__restrict__ a, const V *__restrict__ b,
                                                  int mode, double2 *__restrict__ partial) {
  dot_body<V>(n, B, a, b, mode, partial, blockIdx.x, gridDim.x);
}

// partials of ||v||^2 (x, the same sums in the same order as dot_kernel's
// mode 1) and of the Jacobi-scaled ||D^-1/2 v||^2 (y) per frequency, D the
// l1 row bound of A(w_f) with the operator's coefficients coefa
template <typename T, typename V>
__device__ inline void wnorm_body(int64_t n, int B, const V *__restrict__ v, const OpArgs<T> &la,
                                  const OpCoef *__restrict__ coefa, double2 *__restrict__ partial, int blk,
                                  int nblocks) {
  const int t = threadIdx.x, f = t % B, rl = t / B, R = 256 / B;
  const int64_t chunk = (n + nblocks - 1) / nblocks;
  const int64_t r0 = blk * chunk, r1 = min(n, r0 + chunk);
  double2 acc = {0, 0};
  const CoefL1 ca = coef_l1(coefa[f]);  // the column's bound coefficients, once
  for (int64_t r = r0 + rl; r < r1; r += R) {
    const V a = v[r * B + f];
    const RowL1 w = row_l1<T>(la, r);
    const double q = (double)a.x * a.x + (double)a.y * a.y;
    acc.x += q;
    acc.y += q * row_weight(w, ca);
  }
  red_block_store(acc, B, partial, blk);
}

template <typename T, typename V>
__global__ __launch_bounds__(256) void wnorm_kernel(int64_t n, int B, const V *__restrict__ v, OpArgs<T> la,
                                                    const OpCoef *__restrict__ coefa, double2 *__restrict__ partial) {
  wnorm_body<T, V>(n, B, v, la, coefa, partial, blockIdx.x, gridDim.x);
}

// x += alpha p ; r -= alpha q ; partial sums of |r|^2 and of the
// Jacobi-scaled |r|^2 (wnorm_kernel's layout); with u != nullptr also
// u = D^-1 r, the first l1-Jacobi sweep of the preconditioner
// (jacobi0_kernel's values)
template <typename T>
__device__ inline void update_norm_body(int64_t n, int B, const double2 *__restrict__ alpha,
                                        const typename VT<T>::t *__restrict__ p,
                                        const typename VT<T>::t *__restrict__ q, typename VT<T>::t *__restrict__ x,
                                        typename VT<T>::t *__restrict__ r, double2 *__restrict__ partial,
                                        const OpArgs<T> &la, const OpCoef *__restrict__ coefp,
                                        const OpCoef *__restrict__ coefa, typename VT<T>::t *__restrict__ u, int blk,
                                        int nblocks) {
  using V = typename VT<T>::t;
  const int t = threadIdx.x, f = t % B, rl = t / B, R = 256 / B;
  const int64_t chunk = (n + nblocks - 1) / nblocks;
  const int64_t r0 = blk * chunk, r1 = min(n, r0 + chunk);
  const double2 a = alpha[f];
  const bool upd = !(a.x == 0 && a.y == 0);  // frozen column: x, r untouched
  // the column's bound coefficients, once; every load of a row is issued
  // before its first use (stores count against the same wait counter as
  // loads here, so a load after a store waits for the store too)
  const CoefL1 cp = coef_l1(coefp[f]), ca = coef_l1(coefa[f]);
  double2 acc = {0, 0};
  // the next row's loads are issued before this row's stores
  struct In {
    V rv, pv, qv, xv;
    RowL1 w;
  };
  auto load = [&](int64_t row) {
    const int64_t i = row * B + f;
    In t;
    t.rv = r[i];
    t.pv = t.qv = t.xv = V{};
    if (upd) {
      t.pv = p[i];
      t.qv = q[i];
      t.xv = x[i];
    }
    t.w = row_l1<T>(la, row);
    return t;
  };
  int64_t row = r0 + rl;
  In nx{};
  if (row < r1) nx = load(row);
  for (; row < r1; row += R) {
    const int64_t i = row * B + f;
    const In c = nx;
    if (row + R < r1) nx = load(row + R);
    V rv = c.rv, xv = c.xv;
    const V pv = c.pv, qv = c.qv;
    const RowL1 &w = c.w;
    if (upd) {
      xv.x += a.x * pv.x - a.y * pv.y;
      xv.y += a.x * pv.y + a.y * pv.x;
      rv.x -= a.x * qv.x - a.y * qv.y;
      rv.y -= a.x * qv.y + a.y * qv.x;
    }
    const V uv = jacobi0_value<T>(w, cp, rv);
    const double q2 = (double)rv.x * rv.x + (double)rv.y * rv.y;
    acc.x += q2;
    acc.y += q2 * row_weight(w, ca);
    if (upd) {
      x[i] = xv;
      r[i] = rv;
    }
    if (u) u[i] = uv;
  }
  red_block_store(acc, B, partial, blk);
}

template <typename T>
__global__ __launch_bounds__(256) void cocg_update_norm_kernel(int64_t n, int B, const double2 *__restrict__ alpha,
                                                               const typename VT<T>::t *__restrict__ p,
                                                               const typename VT<T>::t *__restrict__ q,
                                                               typename VT<T>::t *__restrict__ x,
                                                               typename VT<T>::t *__restrict__ r,
                                                               double2 *__restrict__ partial, OpArgs<T> la,
                                                               const OpCoef *__restrict__ coefp,
                                                               const OpCoef *__restrict__ coefa,
                                                               typename VT<T>::t *__restrict__ u) {
  update_norm_body<T>(n, B, alpha, p, q, x, r, partial, la, coefp, coefa, u, blockIdx.x, gridDim.x);
}

__device__ inline void dot_final_body(int nblk, int B, const double2 *__restrict__ partial,
                                      double2 *__restrict__ out) {
  __shared__ double2 sh[kRedFinalThreads];
  const int t = threadIdx.x, f = t % B, rl = t / B, R = kRedFinalThreads / B;
  double2 acc = {0, 0};
  int k = rl;
  for (; k + 7 * R < nblk; k += 8 * R) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = partial[(k + u * R) * B + f];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      acc.x += v[u].x;
      acc.y += v[u].y;
    }
  }
  for (; k < nblk; k += R) {
    acc.x += partial[k * B + f].x;
    acc.y += partial[k * B + f].y;
  }
  sh[t] = acc;
  __syncthreads();
  for (int st = R / 2; st > 0; st >>= 1) {
    if (rl < st) {
      sh[t].x += sh[t + st * B].x;
      sh[t].y += sh[t + st * B].y;
    }
    __syncthreads();
  }
  if (rl == 0) out[f] = sh[t];
}

__global__ __launch_bounds__(kRedFinalThreads) void dot_final_kernel(int nblk, int B,
                                                                     const double2 *__restrict__ partial,
                                                                     double2 *__restrict__ out) {
  dot_final_body(nblk, B, partial, out);
}

// COCG scalars on the device: alpha = rho / mu (0 if inactive), beta = rho'/rho.
// active[f]: 0 = frozen column, 1 = iterating, 2 = (re)starting (beta = 0, p = z)
__device__ inline void alpha_body(int B, const double2 *rho, const double2 *mu, const int *active, double2 *alpha) {
  const int f = threadIdx.x;
  if (f >= B) return;
  double2 a = {0, 0};
  const double den = mu[f].x * mu[f].x + mu[f].y * mu[f].y;
  if (active[f] && den > 0 && isfinite(den)) {
    a.x = (rho[f].x * mu[f].x + rho[f].y * mu[f].y) / den;
    a.y = (rho[f].y * mu[f].x - rho[f].x * mu[f].y) / den;
  }
  alpha[f] = a;
}
__global__ void cocg_alpha_kernel(int B, const double2 *rho, const double2 *mu, const int *active, double2 *alpha) {
  alpha_body(B, rho, mu, active, alpha);
}

__device__ inline void beta_body(int B, const double2 *rho_new, double2 *rho, const int *active, double2 *beta) {
  const int f = threadIdx.x;
  if (f >= B) return;
  double2 b = {0, 0};
  const double den = rho[f].x * rho[f].x + rho[f].y * rho[f].y;
  if (active[f] == 1 && den > 0 && isfinite(den)) {
    b.x = (rho_new[f].x * rho[f].x + rho_new[f].y * rho[f].y) / den;
    b.y = (rho_new[f].y * rho[f].x - rho_new[f].x * rho[f].y) / den;
  }
  beta[f] = b;
  rho[f] = rho_new[f];
}
__global__ void cocg_beta_kernel(int B, const double2 *rho_new, double2 *rho, const int *active, double2 *beta) {
  beta_body(B, rho_new, rho, active, beta);
}

// x = 0 on the rows whose field is not `keep`
template <typename V>
__global__ void keep_field_kernel(int64_t nt, int B, const uint8_t *__restrict__ field, int keep, V *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt || field[i / B] == keep) return;
  x[i] = V{0, 0};
}

// after p = z + beta p: columns that (re)started (2) iterate from now on (1)
__global__ void clear_restart_kernel(int B, int *active) {
  const int k = threadIdx.x;
  if (k < B && active[k] == 2) active[k] = 1;
}

// p = z + beta p
template <typename V>
__device__ inline void dir_body(int64_t nt, int B, const double2 *__restrict__ beta, const V *__restrict__ z,
                                V *__restrict__ p, int64_t i) {
  if (i >= nt) return;
  const double2 bt = beta[i % B];
  const V zv = z[i];
  if (bt.x == 0 && bt.y == 0) {  // (re)start: p = z, whatever p held
    p[i] = zv;
    return;
  }
  V pv = p[i];
  V o;
  o.x = zv.x + bt.x * pv.x - bt.y * pv.y;
  o.y = zv.y + bt.x * pv.y + bt.y * pv.x;
  p[i] = o;
}
template <typename V>
__global__ void cocg_dir_kernel(int64_t nt, int B, const double2 *__restrict__ beta, const V *__restrict__ z,
                                V *__restrict__ p) {
  dir_body<V>(nt, B, beta, z, p, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// refinement: r_in = r64 * s_f (to the inner precision), and x64 += d * w_f
// (perm: row order of the inner operator, nullptr = same order)
template <typename V>
__global__ void scale_to_kernel(int64_t nt, int B, const double *__restrict__ s, const double2 *__restrict__ r64,
                                V *__restrict__ out, const int32_t *__restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const int64_t r = i / B, f = i % B;
  const double sc = s[f];
  const double2 v = r64[(perm ? perm[r] : r) * B + f];
  V o;
  o.x = v.x * sc;
  o.y = v.y * sc;
  out[i] = o;
}

template <typename V>
__global__ void add_scaled_kernel(int64_t nt, int B, const double *__restrict__ w, const V *__restrict__ d,
                                  double2 *__restrict__ x, const int32_t *__restrict__ iperm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const int64_t r = i / B, f = i % B;
  const double sc = w[f];
  if (sc == 0.0) return;
  const V dv = d[(iperm ? iperm[r] : r) * B + f];
  double2 v = x[i];
  v.x += sc * (double)dv.x;
  v.y += sc * (double)dv.y;
  x[i] = v;
}

__global__ void bcast_rhs_kernel(int64_t n, int B, const double2 *__restrict__ F, double2 *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  b[i] = F[i / B];
}

__global__ void unpack_kernel(int64_t n, int B, int nb, const double2 *__restrict__ x, double2 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nb) return;
  const int64_t f = i / n, r = i % n;
  out[i] = x[r * B + f];
}

// slot scheduler: column c of the [row][B] layout to a contiguous vector
__global__ void unpack_col_kernel(int64_t n, int B, int c, const double2 *__restrict__ x, double2 *__restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  out[r] = x[r * B + c];
}

// FP64 refinement of a few frozen columns: add_scaled_kernel's update
// x64[r][map[c]] += w[map[c]] * d[iperm[r]][map[c]] on the nfz frozen
// columns c < nfz (w = 0: unchanged), with the result also gathered into a
// narrow [row][CB] block xc[r][c]; the padding columns c >= nfz of xc are 0.
// narrow_residual_kernel then stores r64[r][map[c]] = b64[r][map[c]] -
// (A xc)[r][c] for c < nfz (the same values as b - A x computed on all B
// columns: each column is summed alone)
template <typename V>
__global__ void add_gather_kernel(int64_t n, int B, int CB, int nfz, const int *__restrict__ map,
                                  const double *__restrict__ w, const V *__restrict__ d,
                                  const int32_t *__restrict__ iperm, double2 *__restrict__ x,
                                  double2 *__restrict__ xc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * CB) return;
  const int c = (int)(i % CB);
  if (c >= nfz) {
    xc[i] = {0, 0};
    return;
  }
  const int64_t r = i / CB;
  const int f = map[c];
  const int64_t o = r * B + f;
  double2 v = x[o];
  const double sc = w[f];
  if (sc != 0.0) {
    const V dv = d[(iperm ? iperm[r] : r) * B + f];
    v.x += sc * (double)dv.x;
    v.y += sc * (double)dv.y;
    x[o] = v;
  }
  xc[i] = v;
}

// columns taking a new frequency: x64 = g (the prescribed values on
// Dirichlet rows, 0 elsewhere; g = nullptr: all 0) and r64 = b - A x64,
// which is b with the Dirichlet rows zeroed (identity rows, zeroed columns)
__global__ void slot_fresh_kernel(int64_t n, int B, const int *__restrict__ fresh, const double2 *__restrict__ b64,
                                  const double2 *__restrict__ g, const uint8_t *__restrict__ field,
                                  double2 *__restrict__ x64, double2 *__restrict__ r64) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int f = (int)(i % B);
  if (!fresh[f]) return;
  const int64_t r = i / B;
  if (g) {
    x64[i] = g[r];
    r64[i] = field[r] == 2 ? double2{0, 0} : b64[i];
  } else {
    x64[i] = {0, 0};
    r64[i] = b64[i];
  }
}

// Prescribed nonzero values (MatZeroRows + F[row] = value,
// analysis.cpp:419-454) lifted into the right-hand side of a fresh column f:
// b[r][f] = F[r] - sum over Dirichlet columns c of A(w_f)[r][c] g[c] on
// the free rows r that touch a Dirichlet column (lift_rows); the solver's
// operator has those columns zeroed, so it stays complex symmetric.  A uses
// the model's unzeroed K, M and extra entries with the frequency's
// coefficients, as the operator does.
__global__ void lift_rhs_kernel(int64_t nl, const int32_t *__restrict__ lift_rows, int B,
                                const int *__restrict__ fresh, const int32_t *__restrict__ rowptr,
                                const int32_t *__restrict__ colind, const uint8_t *__restrict__ field,
                                const double *__restrict__ K, const double *__restrict__ M,
                                const int32_t *__restrict__ xrp, const int32_t *__restrict__ xcol,
                                const uint8_t *__restrict__ xtype, const double *__restrict__ xval,
                                const double2 *__restrict__ g, const OpCoef *__restrict__ coef,
                                const double2 *__restrict__ F, double2 *__restrict__ b64) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl * B) return;
  const int f = (int)(i % B);
  if (!fresh[f]) return;
  const int32_t r = lift_rows[i / B];
  const OpCoef &cf = coef[f];
  const bool fl = field[r] == 1;
  const double2 ck = fl ? cf.kf : cf.ks, cm = fl ? cf.mf : cf.ms;
  double2 acc = {0, 0};
  for (int32_t p = rowptr[r]; p < rowptr[r + 1]; p++) {
    const int32_t c = colind[p];
    if (field[c] != 2 || c == r) continue;
    const double2 a = {ck.x * K[p] + cm.x * M[p], ck.y * K[p] + cm.y * M[p]};
    const double2 t = cmul(a, g[c]);
    acc.x += t.x;
    acc.y += t.y;
  }
  for (int32_t k = xrp[r]; k < xrp[r + 1]; k++) {
    const int32_t c = xcol[k];
    if (field[c] != 2) continue;
    const double2 a = xtype[k] == 0 ? double2{cf.g.x * xval[k], cf.g.y * xval[k]} : double2{xval[k], 0.0};
    const double2 t = cmul(a, g[c]);
    acc.x += t.x;
    acc.y += t.y;
  }
  b64[(int64_t)r * B + f] = {F[r].x - acc.x, F[r].y - acc.y};
}

// columns (re)starting the inner solve: x_in = 0, r_in = r64 * s_f
template <typename T>
__global__ void slot_restart_kernel(int64_t nt, int B, const int *__restrict__ active, const double *__restrict__ s,
                                    const double2 *__restrict__ r64, typename VT<T>::t *__restrict__ r,
                                    typename VT<T>::t *__restrict__ x, const int32_t *__restrict__ perm,
                                    OpArgs<T> la, const OpCoef *__restrict__ coefp, typename VT<T>::t *__restrict__ u) {
  using V = typename VT<T>::t;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const int64_t row = i / B, f = i % B;
  if (active[f] != 2) return;
  const double sc = s[f];
  const double2 v = r64[(perm ? perm[row] : row) * B + f];
  V o;
  o.x = v.x * sc;
  o.y = v.y * sc;
  r[i] = o;
  x[i] = {0, 0};
  if (u) u[i] = jacobi0_value<T>(la, coefp, row, (int)f, o);
}

// solver shadows: FP32 copies, Dirichlet columns zeroed; row l1 sums
template <typename T>
__global__ void shadow_kernel(int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
                              const uint8_t *__restrict__ field, const double *__restrict__ K,
                              const double *__restrict__ M, T *__restrict__ Ko, T *__restrict__ Mo,
                              double *__restrict__ rK, double *__restrict__ rM) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double sk = 0, sm = 0;
  for (int32_t p = rowptr[r]; p < rowptr[r + 1]; p++) {
    const bool z = field[colind[p]] == 2 && colind[p] != r;
    const double k = z ? 0.0 : K[p], m = z ? 0.0 : M[p];
    if (Ko) {
      Ko[p] = (T)k;
      Mo[p] = (T)m;
    }
    sk += fabs(k);
    sm += fabs(m);
  }
  rK[r] = sk;
  rM[r] = sm;
}

template <typename T>
__global__ void extra_shadow_kernel(int64_t n, const int32_t *__restrict__ xrp, const int32_t *__restrict__ xcol,
                                    const uint8_t *__restrict__ xtype, const uint8_t *__restrict__ field,
                                    const double *__restrict__ xv, T *__restrict__ out, double *__restrict__ rG) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double sg = 0;
  for (int32_t k = xrp[r]; k < xrp[r + 1]; k++) {
    const double v = field[xcol[k]] == 2 ? 0.0 : xv[k];
    out[k] = (T)v;
    if (xtype[k] == 0) sg += fabs(v);
  }
  rG[r] = sg;
}

static int nblk(int64_t n, int per) { return (int)((n + per - 1) / per); }

// relative residual of an inner solve: the larger of the plain and the
// Jacobi-scaled reductions, from ||r||^2, ||D^-1/2 r||^2 (h) and the norms
// of its right-hand side (n0)
static double inner_rel(double2 h, double2 n0) {
  const double a = n0.x > 0 ? std::sqrt(h.x) / n0.x : 0.0;
  const double b = n0.y > 0 ? std::sqrt(h.y) / n0.y : 0.0;
  return std::max(a, b);
}

// ---------------------------------------------------------------------------
// solver state: levels of the hierarchy
// ---------------------------------------------------------------------------
struct LevelData {
  Model *model = nullptr;
  std::unique_ptr<Model> own;
  int64_t n = 0, nnz = 0;
  DArr<float> K32, M32, X32;
  DArr<double> K64s, M64s, X64s;
  const double *K64 = nullptr, *M64 = nullptr, *X64 = nullptr;
  DArr<double> rK, rM, rG;
  // interpolation from the next coarser level
  int64_t nc = 0;
  DArr<int32_t> P_rp, P_ci, R_rp, R_ci;
  DArr<float> P32, R32;
  DArr<double> P64, R64;
  // The FP32 operator may run on a locality-improving row order (rows of a
  // node stay together, nodes in Morton order of their coordinates):
  // perm[new] = old, iperm[old] = new; FP64 structures keep elPaSo's order.
  bool permuted = false;
  DArr<int32_t> perm, iperm, rp32, ci32, xrp32, xcol32, P_rp32, P_ci32, R_ci32;
  DArr<uint8_t> xtype32, field32;
  DArr<double> rK32, rM32, rG32;
  // row-padded FP32 operator (row_op_kernel, 32 frequencies)
  DArr<int32_t> pad_rp;
  DArr<uint32_t> pad_off;
  DArr<float> pad_K, pad_M;
  PadCsr padded() const { return {pad_rp.p, pad_off.p, pad_K.p, pad_M.p}; }
  // row tiles of the padded operator (tile_op_kernel); n_tiles = 0: none
  DArr<int32_t> t_row, t_cptr, t_col;
  DArr<uint32_t> t_loff;
  int64_t n_tiles = 0;
  int t_rows = 0, t_lines = 0;  // rows per tile (before splits), most lines of a tile
  TileCsr tiles() const { return {t_row.p, t_cptr.p, t_col.p, t_loff.p, n_tiles}; }
  // rows grouped g_R to a wavefront (rgrp_op_kernel), g_R = 0: none; the
  // padded copy is then padded per group (pad_R = g_R), else to kRowPad
  DArr<int32_t> g_rows;
  int64_t g_waves = 0;
  int g_R = 0, pad_R = 0;
  // with tiles and groups (tgrp_op_kernel): each tile's first group;
  // t_half: each tile's frequencies split over two workgroups (thgrp_op_kernel)
  DArr<int32_t> t_grp;
  bool t_half = false;
  // box cycle: K P, M P, C_Gamma P (row-padded, permuted rows, lattice
  // columns as byte offsets) and P rows padded to 8 (box_post_kernel)
  DArr<int32_t> ap_rp;
  DArr<uint32_t> ap_off, pp_off;
  DArr<float> ap_K, ap_M, ap_G, pp_v;
  PadCsr3 ap() const { return {ap_rp.p, ap_off.p, ap_K.p, ap_M.p, ap_G.p}; }
  // the same on grouped half-line tiles (thbox_post_kernel; bt_on: chosen)
  DArr<int32_t> bp_rp, bt_row, bt_cptr, bt_col, bt_grp, bg_rows;
  DArr<uint32_t> bp_loff;
  DArr<float> bp_K, bp_M, bp_G, bp_P;
  int64_t n_btiles = 0;
  int bt_lines = 0, bt_rows = 0;  // most lines of a tile, rows per tile (before splits)
  bool bt_on = false;
  BoxTiles box_tiles() const {
    return {bp_rp.p, bp_loff.p, bp_K.p, bp_M.p, bp_G.p, bp_P.p, (int64_t)bp_K.n,
            TileCsr{bt_row.p, bt_cptr.p, bt_col.p, nullptr, n_btiles}, bt_grp.p, bg_rows.p};
  }
  template <typename T> OpArgs<T> args() const;
  struct Csr {
    const int32_t *rp, *ci;
  };
  template <typename T> Csr Pcsr() const;
  template <typename T> Csr Rcsr() const;
};

template <>
OpArgs<float> LevelData::args<float>() const {
  Model &m = *model;
  if (permuted)
    return {n, rp32.p, ci32.p, K32.p, M32.p, xrp32.p, xcol32.p, xtype32.p, X32.p, field32.p, rK32.p, rM32.p, rG32.p};
  return {n, m.d_rowptr.p, m.d_colind.p, K32.p, M32.p, m.d_x_rowptr.p, m.d_x_col.p, m.d_x_type.p, X32.p,
          m.d_row_field.p, rK.p, rM.p, rG.p};
}
template <> LevelData::Csr LevelData::Pcsr<float>() const {
  return permuted ? Csr{P_rp32.p, P_ci32.p} : Csr{P_rp.p, P_ci.p};
}
template <> LevelData::Csr LevelData::Pcsr<double>() const { return Csr{P_rp.p, P_ci.p}; }
template <> LevelData::Csr LevelData::Rcsr<float>() const {
  return permuted ? Csr{R_rp.p, R_ci32.p} : Csr{R_rp.p, R_ci.p};
}
template <> LevelData::Csr LevelData::Rcsr<double>() const { return Csr{R_rp.p, R_ci.p}; }
template <>
OpArgs<double> LevelData::args<double>() const {
  Model &m = *model;
  return {n, m.d_rowptr.p, m.d_colind.p, K64, M64, m.d_x_rowptr.p, m.d_x_col.p, m.d_x_type.p, X64,
          m.d_row_field.p, rK.p, rM.p, rG.p};
}

}  // namespace elp

struct elp::Model::SolverState {
  std::vector<std::unique_ptr<LevelData>> levels;
  bool mg = false;
  int mg_levels = 0, mg_min_nodes = 0;
  // exact separable Hex8 solve on box meshes (used instead of the
  // rediscretized hierarchy when available)
  std::unique_ptr<elp::FdmBox, void (*)(elp::FdmBox *)> fdm{nullptr, nullptr};
  int order = 1;
  // coupled models: the box solve on the fluid rows and the direct solve of
  // the structure block
  bool coupled = false, cpl_requested = false;
  std::unique_ptr<elp::BcrPlan, void (*)(elp::BcrPlan *)> bcr{nullptr, nullptr};
  // models with structure and fluid rows iterate in FP64 (see sweep)
  bool fp64_inner = false;
  // the box of the box solve: the model's, or for a fluid mesh that is no
  // box a stand-in lattice over its bounding box (standin = true)
  BoxInfo box;
  bool standin = false;
  int precond = -1;  // requested (the cache key)
  // the 32-slot FP32 operator's variant, chosen at the first sweep that
  // uses it (op_chosen) for the requested op_tile_rows
  int op_tile_rows = 0;
  bool op_chosen = false;
  bool box_chosen = false;  // the post-smoothing kernel was chosen (thbox or box_post)
  // the last sweep's solver of a small model (work vectors, captured
  // iteration graph), kept for the next sweep with the same options;
  // declared last: destroyed before the levels it refers to
  std::unique_ptr<void, void (*)(void *)> kept_solver{nullptr, nullptr};
  int kept_kind = -1;
  elp_solve_opts kept_opts{};
};

namespace elp {

void default_opts(elp_solve_opts &o) {
  o.precision = ELP_PREC_MIXED;
  o.method = ELP_METHOD_COCG;
  o.precond = ELP_PC_MG;
  o.block = 32;
  o.rtol = 1e-8;
  o.max_iters = 4000;
  o.inner_rtol = 3e-3;
  o.shift = -1.0;  // automatic: 0 with the exact box coarse solve, 0.5 otherwise
  o.smooth = 1;
  o.coarse_sweeps = 20;
  o.mg_levels = 0;
  o.mg_min_nodes = 2000;
  o.op_tile_rows = 0;
  o.graph = 0;
}

static void prepare_level(LevelData &L, hipStream_t s) {
  Model &m = *L.model;
  Ctx *c = m.ctx;
  L.n = m.n;
  L.nnz = m.nnz;
  L.K32.alloc(c, m.nnz);
  L.M32.alloc(c, m.nnz);
  L.rK.alloc(c, m.n);
  L.rM.alloc(c, m.n);
  L.rG.alloc(c, m.n);
  const bool dir = !m.dir_rows.empty();
  hipLaunchKernelGGL(shadow_kernel<float>, dim3(nblk(m.n, 256)), dim3(256), 0, s, m.n, m.d_rowptr.p, m.d_colind.p,
                     m.d_row_field.p, m.d_K.p, m.d_M.p, L.K32.p, L.M32.p, L.rK.p, L.rM.p);
  ELP_HIP(hipGetLastError());
  if (dir) {
    L.K64s.alloc(c, m.nnz);
    L.M64s.alloc(c, m.nnz);
    hipLaunchKernelGGL(shadow_kernel<double>, dim3(nblk(m.n, 256)), dim3(256), 0, s, m.n, m.d_rowptr.p,
                       m.d_colind.p, m.d_row_field.p, m.d_K.p, m.d_M.p, L.K64s.p, L.M64s.p, L.rK.p, L.rM.p);
    ELP_HIP(hipGetLastError());
    L.K64 = L.K64s.p;
    L.M64 = L.M64s.p;
  } else {
    L.K64 = m.d_K.p;
    L.M64 = m.d_M.p;
  }
  const int64_t nx = (int64_t)m.x_col.size();
  L.X32.alloc(c, std::max<int64_t>(nx, 1));
  L.X64s.alloc(c, std::max<int64_t>(nx, 1));
  hipLaunchKernelGGL(extra_shadow_kernel<float>, dim3(nblk(m.n, 256)), dim3(256), 0, s, m.n, m.d_x_rowptr.p,
                     m.d_x_col.p, m.d_x_type.p, m.d_row_field.p, m.d_x_val.p, L.X32.p, L.rG.p);
  hipLaunchKernelGGL(extra_shadow_kernel<double>, dim3(nblk(m.n, 256)), dim3(256), 0, s, m.n, m.d_x_rowptr.p,
                     m.d_x_col.p, m.d_x_type.p, m.d_row_field.p, m.d_x_val.p, L.X64s.p, L.rG.p);
  ELP_HIP(hipGetLastError());
  L.X64 = L.X64s.p;
}

// Hex8 fluid box with consistent Robin faces on the requested sides
static std::unique_ptr<Model> make_box_level(const Model &fine, const int nel[3], std::complex<double> Z) {
  auto m = std::make_unique<Model>();
  m->ctx_ref = fine.ctx_ref;
  m->ctx = fine.ctx;
  const int px = nel[0] + 1, py = nel[1] + 1, pz = nel[2] + 1;
  const int64_t nn = (int64_t)px * py * pz;
  m->node_ids.resize(nn);
  m->xyz.resize(nn * 3);
  for (int64_t k = 0; k < nn; k++) {
    int64_t i = k % px, j = (k / px) % py, l = k / ((int64_t)px * py);
    m->node_ids[k] = k + 1;
    m->xyz[3 * k] = fine.box.origin[0] + fine.box.length[0] * i / nel[0];
    m->xyz[3 * k + 1] = fine.box.origin[1] + fine.box.length[1] * j / nel[1];
    m->xyz[3 * k + 2] = fine.box.origin[2] + fine.box.length[2] * l / nel[2];
  }
  Material mat = fine.mats.begin()->second;
  m->mats[1] = mat;
  ElemGroupIn g;
  g.kind = ELP_FLUID8;
  g.mat = 1;
  static const int nat[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const int64_t ne = (int64_t)nel[0] * nel[1] * nel[2];
  for (int64_t e = 0; e < ne; e++) {
    int64_t i = e % nel[0], j = (e / nel[0]) % nel[1], l = e / ((int64_t)nel[0] * nel[1]);
    g.ids.push_back(e + 1);
    for (int a = 0; a < 8; a++)
      g.conn.push_back((i + nat[a][0]) + px * ((j + nat[a][1]) + (int64_t)py * (l + nat[a][2])) + 1);
    // Robin sides: Hex8 face f lies on the box side f when the element touches it
    const bool on[6] = {j == 0, j == nel[1] - 1, l == 0, i == nel[0] - 1, l == nel[2] - 1, i == 0};
    for (int f = 0; f < 6; f++)
      if ((fine.box.robin_sides >> f & 1) && on[f]) {
        m->imp_elem.push_back(e + 1);
        m->imp_face.push_back(f + 1);
      }
  }
  m->groups.push_back(std::move(g));
  m->Z = Z;
  m->imp_lumped = false;
  return m;
}

// ---------------------------------------------------------------------------
// locality-improving row order for the FP32 operator
// ---------------------------------------------------------------------------
static uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

template <typename T>
__global__ void permute_rows_kernel(int64_t n, const int32_t *__restrict__ perm, const int32_t *__restrict__ iperm,
                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                    const T *__restrict__ a, const T *__restrict__ b, const int32_t *__restrict__ rpn,
                                    int32_t *__restrict__ cin, T *__restrict__ an, T *__restrict__ bn) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t o = perm[r];
  int32_t q = rpn[r];
  for (int32_t p = rp[o]; p < rp[o + 1]; p++, q++) {
    cin[q] = iperm[ci[p]];
    an[q] = a[p];
    if (b) bn[q] = b[p];
  }
}

// row-padded copy for row_op_kernel (one thread per row)
__global__ void pad_rows_kernel(int64_t n, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                const float *__restrict__ K, const float *__restrict__ M,
                                const int32_t *__restrict__ rp8, uint32_t *__restrict__ off8, float *__restrict__ K8,
                                float *__restrict__ M8, uint32_t line_bytes, uint32_t oob) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int32_t q = rp8[r];
  for (int32_t p = rp[r]; p < rp[r + 1]; p++, q++) {
    off8[q] = (uint32_t)ci[p] * line_bytes;
    K8[q] = K[p];
    M8[q] = M[p];
  }
  for (; q < rp8[r + 1]; q++) {
    off8[q] = oob;
    K8[q] = 0.f;
    M8[q] = 0.f;
  }
}

template <typename T>
__global__ void gather_perm_kernel(int64_t n, const int32_t *__restrict__ perm, const T *__restrict__ in,
                                   T *__restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) out[r] = in[perm[r]];
}

// K P, M P and C_Gamma P for box_post_kernel, in the FP32 operator's row
// order, from the FP32 values the operator itself applies (so that both
// smoothing steps of the cycle see the same operator): the products of FP32
// entries are summed in FP64 over the row's entries in row order and rounded
// once.  Built on the host in parallel over row ranges; skipped (the cycle
// then uses the plain operator) when a model has FSI entries or exceeds the
// 32-bit offsets.
static void build_box_products(LevelData &L, hipStream_t s) {
  Model &m = *L.model;
  const int64_t n = m.n, nc = L.nc;
  if (n > kRowOpMaxRows || nc <= 0 || nc > kRowOpMaxRows || !L.permuted) return;
  for (auto t : m.x_type)
    if (t != 0) return;
  constexpr uint32_t line = 32 * sizeof(float2);
  ELP_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> rp, ci, xrp, xci, prp, pci;
  std::vector<float> K, M, X, pv;
  L.rp32.download(rp);
  L.ci32.download(ci);
  L.K32.download(K);
  L.M32.download(M);
  L.xrp32.download(xrp);
  L.xcol32.download(xci);
  L.X32.download(X);
  L.P_rp32.download(prp);
  L.P_ci32.download(pci);
  L.P32.download(pv);
  for (int64_t r = 0; r < n; r++)
    if (prp[r + 1] - prp[r] > 8) return;
  struct Part {
    std::vector<int32_t> len;
    std::vector<uint32_t> off;
    std::vector<float> K, M, G;
  };
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int64_t chunk = (n + nt - 1) / nt;
  std::vector<Part> part(nt);
  auto work = [&](unsigned t) {
    const int64_t a = (int64_t)t * chunk, b = std::min<int64_t>(n, a + chunk);
    if (a >= b) return;
    std::vector<int32_t> stamp(nc, -1), list;
    std::vector<double> aK(nc), aM(nc), aG(nc);
    Part &P = part[t];
    for (int64_t i = a; i < b; i++) {
      list.clear();
      auto add = [&](int32_t j, double kv, double mv, double gv) {
        for (int32_t q = prp[j]; q < prp[j + 1]; q++) {
          const int32_t c = pci[q];
          const double w = pv[q];
          if (stamp[c] != (int32_t)i) {
            stamp[c] = (int32_t)i;
            aK[c] = aM[c] = aG[c] = 0.0;
            list.push_back(c);
          }
          aK[c] += kv * w;
          aM[c] += mv * w;
          aG[c] += gv * w;
        }
      };
      for (int32_t p = rp[i]; p < rp[i + 1]; p++) add(ci[p], K[p], M[p], 0.0);
      for (int32_t p = xrp[i]; p < xrp[i + 1]; p++) add(xci[p], 0.0, 0.0, X[p]);
      std::sort(list.begin(), list.end());
      const int len = (int)list.size(), padded = (len + kRowPad - 1) / kRowPad * kRowPad;
      P.len.push_back(padded);
      for (int32_t c : list) {
        P.off.push_back((uint32_t)c * line);
        P.K.push_back((float)aK[c]);
        P.M.push_back((float)aM[c]);
        P.G.push_back((float)aG[c]);
      }
      for (int k = len; k < padded; k++) {
        P.off.push_back((uint32_t)(nc * line));
        P.K.push_back(0.f);
        P.M.push_back(0.f);
        P.G.push_back(0.f);
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++) th.emplace_back(work, t);
  for (auto &x : th) x.join();
  std::vector<int32_t> arp(1, 0);
  std::vector<uint32_t> off;
  std::vector<float> vK, vM, vG;
  for (auto &P : part) {
    for (int32_t l : P.len) arp.push_back(arp.back() + l);
    off.insert(off.end(), P.off.begin(), P.off.end());
    vK.insert(vK.end(), P.K.begin(), P.K.end());
    vM.insert(vM.end(), P.M.begin(), P.M.end());
    vG.insert(vG.end(), P.G.begin(), P.G.end());
    P = Part();
  }
  // prolongation rows, 8 entries each
  std::vector<uint32_t> poff((size_t)n * 8, (uint32_t)(nc * line));
  std::vector<float> pval((size_t)n * 8, 0.f);
  for (int64_t i = 0; i < n; i++)
    for (int32_t q = prp[i]; q < prp[i + 1]; q++) {
      poff[i * 8 + (q - prp[i])] = (uint32_t)pci[q] * line;
      pval[i * 8 + (q - prp[i])] = pv[q];
    }
  Ctx *c = m.ctx;
  L.ap_rp.upload(c, arp);
  L.ap_off.upload(c, off);
  L.ap_K.upload(c, vK);
  L.ap_M.upload(c, vM);
  L.ap_G.upload(c, vG);
  L.pp_off.upload(c, poff);
  L.pp_v.upload(c, pval);
}

// The box products and the P rows merged for thbox_post_kernel: per row its
// (A P) entries in column order with the P value of each column (0 where P
// has none), in tiles of `rows` consecutive rows (halved until their
// distinct lattice columns number at most 255), the rows of a tile taken by
// length 4 at a time and padded to one length per group; per entry the
// LDS offset of its column's 128-byte half-line in the tile.  Returns false
// (nothing built) when a P column is missing from its (A P) row.
static bool build_box_tiles(LevelData &L, int rows) {
  Model &m = *L.model;
  const int64_t n = L.n, nc = L.nc;
  constexpr uint32_t line = 32 * sizeof(float2);
  std::vector<int32_t> arp;
  std::vector<uint32_t> aoff, poff;
  std::vector<float> aK, aM, aG, pv;
  L.ap_rp.download(arp);
  L.ap_off.download(aoff);
  L.ap_K.download(aK);
  L.ap_M.download(aM);
  L.ap_G.download(aG);
  L.pp_off.download(poff);
  L.pp_v.download(pv);
  // merged rows: lattice columns and P values
  std::vector<int32_t> mrp(n + 1, 0), mcol;
  std::vector<float> mP;
  mcol.reserve(aoff.size());
  mP.reserve(aoff.size());
  for (int64_t r = 0; r < n; r++) {
    const int32_t a0 = (int32_t)mcol.size();
    for (int32_t p = arp[r]; p < arp[r + 1]; p++) {
      if (aoff[p] >= (uint32_t)(nc * line)) continue;  // padding
      mcol.push_back((int32_t)(aoff[p] / line));
      mP.push_back(0.f);
    }
    for (int t = 0; t < 8; t++) {
      const uint32_t o = poff[r * 8 + t];
      if (o >= (uint32_t)(nc * line)) continue;
      const int32_t c = (int32_t)(o / line);
      const auto it = std::lower_bound(mcol.begin() + a0, mcol.end(), c);
      if (it == mcol.end() || *it != c) return false;
      mP[it - mcol.begin()] = pv[r * 8 + t];
    }
    mrp[r + 1] = (int32_t)mcol.size();
  }
  // the K, M, G values of the merged entries: the unpadded (A P) entries
  std::vector<float> vK(mcol.size()), vM(mcol.size()), vG(mcol.size());
  for (int64_t r = 0, q = 0; r < n; r++)
    for (int32_t p = arp[r]; p < arp[r + 1]; p++) {
      if (aoff[p] >= (uint32_t)(nc * line)) continue;
      vK[q] = aK[p];
      vM[q] = aM[p];
      vG[q] = aG[p];
      q++;
    }
  // tiles
  struct Sub {
    int32_t r0, r1;
    std::vector<int32_t> cols;
  };
  const int64_t nch = (n + rows - 1) / rows;
  std::vector<std::vector<Sub>> per(nch);
  parallel_for(nch, [&](int64_t a, int64_t b) {
    std::vector<int32_t> u;
    for (int64_t ch = a; ch < b; ch++) {
      std::vector<std::pair<int32_t, int32_t>> todo{{(int32_t)(ch * rows), (int32_t)std::min<int64_t>(n, (ch + 1) * rows)}};
      while (!todo.empty()) {
        const auto rr = todo.back();
        todo.pop_back();
        u.assign(mcol.begin() + mrp[rr.first], mcol.begin() + mrp[rr.second]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if ((int)u.size() > kTileMaxLines && rr.second - rr.first > 1) {
          const int32_t mid = rr.first + (rr.second - rr.first) / 2;
          todo.push_back({mid, rr.second});
          todo.push_back({rr.first, mid});
          continue;
        }
        per[ch].push_back({rr.first, rr.second, u});
      }
    }
  });
  std::vector<const Sub *> tl;
  int most = 0;
  for (auto &v : per)
    for (auto &t : v) {
      tl.push_back(&t);
      most = std::max(most, (int)t.cols.size());
    }
  if (most > kTileMaxLines) return false;
  // groups of 4 rows per tile, padded per group
  std::vector<int32_t> plen(n, 0), idx, tgrp{0}, grows;
  for (const Sub *t : tl) {
    idx.resize(t->r1 - t->r0);
    for (int32_t r = t->r0; r < t->r1; r++) idx[r - t->r0] = r;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int32_t a, int32_t b) { return mrp[a + 1] - mrp[a] > mrp[b + 1] - mrp[b]; });
    for (size_t k = 0; k < idx.size(); k += 4) {
      const int32_t longest = mrp[idx[k] + 1] - mrp[idx[k]];
      const int32_t glen = (longest + kRowPad - 1) / kRowPad * kRowPad;
      for (int u = 0; u < 4; u++) {
        if (k + u < idx.size()) plen[idx[k + u]] = glen;
        grows.push_back(k + u < idx.size() ? idx[k + u] : -1);
      }
    }
    tgrp.push_back((int32_t)(grows.size() / 4));
  }
  std::vector<int32_t> brp(n + 1, 0);
  for (int64_t r = 0; r < n; r++) brp[r + 1] = brp[r] + plen[r];
  const int64_t np = brp[n];
  std::vector<uint32_t> loff(np, 0);
  std::vector<float> bK(np, 0.f), bM(np, 0.f), bG(np, 0.f), bP(np, 0.f);
  parallel_for((int64_t)tl.size(), [&](int64_t a, int64_t b) {
    for (int64_t k = a; k < b; k++) {
      const Sub &t = *tl[k];
      for (int32_t r = t.r0; r < t.r1; r++) {
        int32_t q = brp[r];
        for (int32_t p = mrp[r]; p < mrp[r + 1]; p++, q++) {
          loff[q] = 128u * (uint32_t)(1 + (std::lower_bound(t.cols.begin(), t.cols.end(), mcol[p]) - t.cols.begin()));
          bK[q] = vK[p];
          bM[q] = vM[p];
          bG[q] = vG[p];
          bP[q] = mP[p];
        }
      }
    }
  });
  std::vector<int32_t> trow{0}, tcp{0}, tcol;
  for (const Sub *t : tl) {
    trow.back() = t->r0;
    trow.push_back(t->r1);
    tcol.insert(tcol.end(), t->cols.begin(), t->cols.end());
    tcp.push_back((int32_t)tcol.size());
  }
  Ctx *c = m.ctx;
  L.bp_rp.upload(c, brp);
  L.bp_loff.upload(c, loff);
  L.bp_K.upload(c, bK);
  L.bp_M.upload(c, bM);
  L.bp_G.upload(c, bG);
  L.bp_P.upload(c, bP);
  L.bt_row.upload(c, trow);
  L.bt_cptr.upload(c, tcp);
  L.bt_col.upload(c, tcol);
  L.bt_grp.upload(c, tgrp);
  L.bg_rows.upload(c, grows);
  L.n_btiles = (int64_t)trow.size() - 1;
  L.bt_lines = most;
  return true;
}

static void release_box_tiles(LevelData &L) {
  L.bp_rp.release();
  L.bp_loff.release();
  L.bp_K.release();
  L.bp_M.release();
  L.bp_G.release();
  L.bp_P.release();
  L.bt_row.release();
  L.bt_cptr.release();
  L.bt_col.release();
  L.bt_grp.release();
  L.bg_rows.release();
  L.n_btiles = 0;
  L.bt_on = false;
}

// padded row pointer of the row kernel and the tiles: each row rounded up
// to a multiple of kRowPad entries
static std::vector<int32_t> row_padding(const std::vector<int32_t> &rp) {
  const int64_t n = (int64_t)rp.size() - 1;
  std::vector<int32_t> rp8(n + 1, 0);
  for (int64_t r = 0; r < n; r++) {
    const int32_t len = rp[r + 1] - rp[r];
    rp8[r + 1] = rp8[r] + (len + kRowPad - 1) / kRowPad * kRowPad;
  }
  return rp8;
}

// Rows grouped R to a wavefront for rgrp_op_kernel: in each window of 64
// consecutive rows (operator order) the rows sorted by length, longest
// first (ties in row order), and taken R at a time; every row of a group is
// padded to the group's longest row rounded up to kRowPad entries.  Returns
// the padded row pointer; grows gets each group's rows (-1 past the end).
static std::vector<int32_t> group_padding(const std::vector<int32_t> &rp, int R, std::vector<int32_t> &grows) {
  constexpr int kWin = 64;
  const int64_t n = (int64_t)rp.size() - 1;
  std::vector<int32_t> plen(n, 0), idx;
  grows.clear();
  for (int64_t w0 = 0; w0 < n; w0 += kWin) {
    const int64_t w1 = std::min<int64_t>(n, w0 + kWin);
    idx.resize(w1 - w0);
    for (int64_t r = w0; r < w1; r++) idx[r - w0] = (int32_t)r;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int32_t u, int32_t v) { return rp[u + 1] - rp[u] > rp[v + 1] - rp[v]; });
    for (size_t k = 0; k < idx.size(); k += R) {
      const int32_t longest = rp[idx[k] + 1] - rp[idx[k]];
      const int32_t glen = (longest + kRowPad - 1) / kRowPad * kRowPad;
      for (int u = 0; u < R; u++) {
        if (k + u < idx.size()) {
          plen[idx[k + u]] = glen;
          grows.push_back(idx[k + u]);
        } else {
          grows.push_back(-1);
        }
      }
    }
  }
  std::vector<int32_t> rp8(n + 1, 0);
  for (int64_t r = 0; r < n; r++) rp8[r + 1] = rp8[r] + plen[r];
  return rp8;
}

// (re)build the padded copy of the level's FP32 operator for the padded row
// pointer rp8 (rows padded per group of pad_R rows, or pad_R = 0)
static void pad_operator(LevelData &L, const std::vector<int32_t> &rp8, int pad_R, hipStream_t s) {
  Model &m = *L.model;
  Ctx *c = m.ctx;
  ELP_REQUIRE((int64_t)rp8.size() == m.n + 1, ELP_ESTATE, "padded row pointer size");
  L.pad_rp.upload(c, rp8);
  L.pad_off.alloc(c, rp8[m.n]);
  L.pad_K.alloc(c, rp8[m.n]);
  L.pad_M.alloc(c, rp8[m.n]);
  const uint32_t line = 32 * sizeof(float2);
  hipLaunchKernelGGL(pad_rows_kernel, dim3(nblk(m.n, 256)), dim3(256), 0, s, m.n, L.rp32.p, L.ci32.p, L.K32.p,
                     L.M32.p, L.pad_rp.p, L.pad_off.p, L.pad_K.p, L.pad_M.p, line, (uint32_t)(m.n * line));
  ELP_HIP(hipGetLastError());
  ELP_HIP(hipStreamSynchronize(s));
  L.pad_R = pad_R;
}

static void permute_level_float(LevelData &L, hipStream_t s) {
  Model &m = *L.model;
  Ctx *c = m.ctx;
  // Morton order of the nodes that carry DOFs
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t r = 0; r < m.n_nodes; r++) {
    if (m.node_row[r] < 0) continue;
    const double *x = &m.xyz[(int64_t)m.rank_to_input[r] * 3];
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], x[d]);
      hi[d] = std::max(hi[d], x[d]);
    }
  }
  std::vector<std::pair<uint64_t, int32_t>> key;
  key.reserve(m.n_nodes);
  for (int64_t r = 0; r < m.n_nodes; r++) {
    if (m.node_row[r] < 0) continue;
    const double *x = &m.xyz[(int64_t)m.rank_to_input[r] * 3];
    uint64_t k = 0;
    for (int d = 0; d < 3; d++) {
      const double span = hi[d] > lo[d] ? hi[d] - lo[d] : 1.0;
      const uint64_t q = (uint64_t)std::min(2097151.0, std::floor((x[d] - lo[d]) / span * 2097151.0 + 0.5));
      k |= spread3(q) << d;
    }
    key.push_back({k, (int32_t)r});
  }
  std::stable_sort(key.begin(), key.end());
  std::vector<int32_t> perm, iperm(m.n);
  perm.reserve(m.n);
  for (auto &kv : key)
    for (int d = 0; d < m.node_nd[kv.second]; d++) perm.push_back(m.node_row[kv.second] + d);
  for (int64_t r = 0; r < m.n; r++) iperm[perm[r]] = (int32_t)r;
  std::vector<int32_t> rpn(m.n + 1, 0), xrpn(m.n + 1, 0);
  for (int64_t r = 0; r < m.n; r++) {
    rpn[r + 1] = rpn[r] + (m.rowptr[perm[r] + 1] - m.rowptr[perm[r]]);
    xrpn[r + 1] = xrpn[r] + (m.x_rowptr[perm[r] + 1] - m.x_rowptr[perm[r]]);
  }
  L.perm.upload(c, perm);
  L.iperm.upload(c, iperm);
  L.rp32.upload(c, rpn);
  L.xrp32.upload(c, xrpn);
  L.ci32.alloc(c, m.nnz);
  DArr<float> Kn, Mn;
  Kn.alloc(c, m.nnz);
  Mn.alloc(c, m.nnz);
  const int g = nblk(m.n, 256);
  hipLaunchKernelGGL(permute_rows_kernel<float>, dim3(g), dim3(256), 0, s, m.n, L.perm.p, L.iperm.p, m.d_rowptr.p,
                     m.d_colind.p, L.K32.p, L.M32.p, L.rp32.p, L.ci32.p, Kn.p, Mn.p);
  const int64_t nx = (int64_t)m.x_col.size();
  L.xcol32.alloc(c, std::max<int64_t>(nx, 1));
  DArr<float> Xn;
  Xn.alloc(c, std::max<int64_t>(nx, 1));
  if (nx)
    hipLaunchKernelGGL(permute_rows_kernel<float>, dim3(g), dim3(256), 0, s, m.n, L.perm.p, L.iperm.p,
                       m.d_x_rowptr.p, m.d_x_col.p, L.X32.p, (const float *)nullptr, L.xrp32.p, L.xcol32.p, Xn.p,
                       (float *)nullptr);
  // extra-entry types follow their rows (copy through the same row map)
  std::vector<uint8_t> xt(std::max<int64_t>(nx, 1), 0);
  for (int64_t r = 0; r < m.n; r++) {
    int32_t q = xrpn[r];
    for (int32_t p = m.x_rowptr[perm[r]]; p < m.x_rowptr[perm[r] + 1]; p++) xt[q++] = m.x_type[p];
  }
  L.xtype32.upload(c, xt);
  std::vector<uint8_t> fld(m.n);
  for (int64_t r = 0; r < m.n; r++) fld[r] = m.row_field[perm[r]];
  L.field32.upload(c, fld);
  L.rK32.alloc(c, m.n);
  L.rM32.alloc(c, m.n);
  L.rG32.alloc(c, m.n);
  hipLaunchKernelGGL(gather_perm_kernel<double>, dim3(g), dim3(256), 0, s, m.n, L.perm.p, L.rK.p, L.rK32.p);
  hipLaunchKernelGGL(gather_perm_kernel<double>, dim3(g), dim3(256), 0, s, m.n, L.perm.p, L.rM.p, L.rM32.p);
  hipLaunchKernelGGL(gather_perm_kernel<double>, dim3(g), dim3(256), 0, s, m.n, L.perm.p, L.rG.p, L.rG32.p);
  ELP_HIP(hipGetLastError());
  if (L.nc) {
    // interpolation: permute the rows of P, map the columns of R
    std::vector<int32_t> prp, pci, rci;
    std::vector<float> pv;
    L.P_rp.download(prp);
    L.P_ci.download(pci);
    L.P32.download(pv);
    L.R_ci.download(rci);
    std::vector<int32_t> prpn(m.n + 1, 0), pcin;
    std::vector<float> pvn;
    for (int64_t r = 0; r < m.n; r++) {
      for (int32_t p = prp[perm[r]]; p < prp[perm[r] + 1]; p++) {
        pcin.push_back(pci[p]);
        pvn.push_back(pv[p]);
      }
      prpn[r + 1] = (int32_t)pcin.size();
    }
    for (auto &v : rci) v = iperm[v];
    L.P_rp32.upload(c, prpn);
    L.P_ci32.upload(c, pcin);
    L.P32.upload(c, pvn);
    L.R_ci32.upload(c, rci);
  }
  ELP_HIP(hipStreamSynchronize(s));
  L.K32 = std::move(Kn);
  L.M32 = std::move(Mn);
  L.X32 = std::move(Xn);
  L.permuted = true;
  if (m.n <= kRowOpMaxRows) pad_operator(L, row_padding(rpn), 0, s);
}


// Row tiles over the padded FP32 operator (rows in the operator's order):
// chunks of `rows` consecutive rows, each halved until its distinct columns
// number at most max_lines (a single row always fits: rows hold at most 256
// columns).  Per padded entry the 1-based position of its column in the
// tile's sorted column list times line_bytes (its line's byte offset in LDS:
// 256, or 128 for the half-lines of thgrp_op_kernel), 0 for padding.  R = 0: tiles of tile_op_kernel over the row padding; R = 2 or 4:
// tiles of tgrp_op_kernel, the rows of each tile taken by length (longest
// first, ties in row order) R at a time and padded to one length per group;
// the padded copy is rebuilt for that padding, tgrp gets each tile's first
// group and g_rows each group's rows (-1 past the tile's last row).
static void build_tiles(LevelData &L, const std::vector<int32_t> &perm, const std::vector<int32_t> &iperm,
                        const std::vector<int32_t> &rp, int rows, int max_lines, int R, hipStream_t s,
                        uint32_t line_bytes = 256) {
  const Model &m = *L.model;
  const int64_t n = m.n, nch = (n + rows - 1) / rows;
  struct Sub {
    int32_t r0, r1;
    std::vector<int32_t> cols;
  };
  std::vector<std::vector<Sub>> per(nch);
  parallel_for(nch, [&](int64_t a, int64_t b) {
    std::vector<int32_t> u;
    for (int64_t ch = a; ch < b; ch++) {
      std::vector<std::pair<int32_t, int32_t>> todo{{(int32_t)(ch * rows), (int32_t)std::min<int64_t>(n, (ch + 1) * rows)}};
      while (!todo.empty()) {  // depth first, first half first: tiles stay in row order
        const auto rr = todo.back();
        todo.pop_back();
        u.clear();
        for (int32_t r = rr.first; r < rr.second; r++)
          for (int32_t p = m.rowptr[perm[r]]; p < m.rowptr[perm[r] + 1]; p++) u.push_back(iperm[m.colind[p]]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if ((int)u.size() > max_lines && rr.second - rr.first > 1) {
          const int32_t mid = rr.first + (rr.second - rr.first) / 2;
          todo.push_back({mid, rr.second});
          todo.push_back({rr.first, mid});
          continue;
        }
        per[ch].push_back({rr.first, rr.second, u});
      }
    }
  });
  std::vector<const Sub *> tl;
  for (auto &v : per)
    for (auto &t : v) tl.push_back(&t);
  // padding: per row, or per group of R rows of a tile
  std::vector<int32_t> rp8, tgrp{0}, grows;
  if (R > 0) {
    std::vector<int32_t> plen(n, 0), idx;
    for (const Sub *t : tl) {
      idx.resize(t->r1 - t->r0);
      for (int32_t r = t->r0; r < t->r1; r++) idx[r - t->r0] = r;
      std::stable_sort(idx.begin(), idx.end(),
                       [&](int32_t u, int32_t v) { return rp[u + 1] - rp[u] > rp[v + 1] - rp[v]; });
      for (size_t k = 0; k < idx.size(); k += R) {
        const int32_t longest = rp[idx[k] + 1] - rp[idx[k]];
        const int32_t glen = (longest + kRowPad - 1) / kRowPad * kRowPad;
        for (int u = 0; u < R; u++) {
          if (k + u < idx.size()) plen[idx[k + u]] = glen;
          grows.push_back(k + u < idx.size() ? idx[k + u] : -1);
        }
      }
      tgrp.push_back((int32_t)(grows.size() / R));
    }
    rp8.assign(n + 1, 0);
    for (int64_t r = 0; r < n; r++) rp8[r + 1] = rp8[r] + plen[r];
  } else {
    rp8 = row_padding(rp);
  }
  std::vector<uint32_t> loff(rp8[n], 0);
  parallel_for((int64_t)tl.size(), [&](int64_t a, int64_t b) {
    for (int64_t k = a; k < b; k++) {
      const Sub &t = *tl[k];
      for (int32_t r = t.r0; r < t.r1; r++) {
        int32_t q = rp8[r];
        for (int32_t p = m.rowptr[perm[r]]; p < m.rowptr[perm[r] + 1]; p++, q++)
          loff[q] = line_bytes * (uint32_t)(1 + (std::lower_bound(t.cols.begin(), t.cols.end(),
                                                                   iperm[m.colind[p]]) -
                                                  t.cols.begin()));
      }
    }
  });
  std::vector<int32_t> trow{0}, tcp{0}, tcol;
  int most = 0;
  for (const Sub *t : tl) {
    trow.back() = t->r0;
    trow.push_back(t->r1);
    tcol.insert(tcol.end(), t->cols.begin(), t->cols.end());
    tcp.push_back((int32_t)tcol.size());
    most = std::max(most, (int)t->cols.size());
  }
  Ctx *c = m.ctx;
  if (R > 0 || L.pad_R != 0) pad_operator(L, rp8, R, s);
  L.t_row.upload(c, trow);
  L.t_cptr.upload(c, tcp);
  L.t_col.upload(c, tcol);
  L.t_loff.upload(c, loff);
  L.n_tiles = (int64_t)trow.size() - 1;
  L.t_rows = rows;
  L.t_lines = most;
  if (R > 0) {
    L.t_grp.upload(c, tgrp);
    L.g_rows.upload(c, grows);
    L.g_waves = (int64_t)grows.size() / R;
    L.g_R = R;
  }
  if (most > max_lines) L.n_tiles = 0;  // a row alone has more columns than fit: no tiles
}

// trilinear interpolation from a coarse box lattice to the rows of `fine`
static void build_interp(LevelData &F, const Model &coarse_like_box, const int nelc[3], bool fluid_only = false) {
  const Model &fm = *F.model;
  const BoxInfo &bx = coarse_like_box.box;
  std::vector<int32_t> rp(1, 0), ci;
  std::vector<double> v;
  std::vector<std::vector<std::pair<int32_t, double>>> cols_of(fm.n);
  const int px = nelc[0] + 1, py = nelc[1] + 1;
  for (int64_t r = 0; r < fm.n_nodes; r++) {
    if (fm.node_row[r] < 0) continue;
    if (fluid_only && fm.row_field[fm.node_row[r]] != 1) continue;
    const int64_t in = fm.rank_to_input[r];
    double t[3];
    int cell[3];
    for (int d = 0; d < 3; d++) {
      const double h = bx.length[d] / nelc[d];
      double s = (fm.xyz[in * 3 + d] - bx.origin[d]) / h;
      int c = (int)std::floor(s);
      c = std::min(std::max(c, 0), nelc[d] - 1);
      t[d] = std::min(std::max(s - c, 0.0), 1.0);
      cell[d] = c;
    }
    auto &lst = cols_of[fm.node_row[r]];
    for (int k = 0; k < 2; k++)
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
          double w = (i ? t[0] : 1 - t[0]) * (j ? t[1] : 1 - t[1]) * (k ? t[2] : 1 - t[2]);
          if (std::fabs(w) < 1e-13) continue;
          int32_t col = (int32_t)((cell[0] + i) + (int64_t)px * ((cell[1] + j) + (int64_t)py * (cell[2] + k)));
          lst.push_back({col, w});
        }
  }
  for (int64_t r = 0; r < fm.n; r++) {
    auto &lst = cols_of[r];
    std::sort(lst.begin(), lst.end());
    for (auto &p : lst) {
      ci.push_back(p.first);
      v.push_back(p.second);
    }
    rp.push_back((int32_t)ci.size());
  }
  const int64_t nc = (int64_t)px * py * (nelc[2] + 1);
  // transpose
  std::vector<int32_t> trp(nc + 1, 0), tci(ci.size());
  std::vector<double> tv(ci.size());
  for (int32_t c : ci) trp[c + 1]++;
  for (int64_t c = 0; c < nc; c++) trp[c + 1] += trp[c];
  std::vector<int32_t> fill(trp.begin(), trp.end() - 1);
  for (int64_t r = 0; r < fm.n; r++)
    for (int32_t p = rp[r]; p < rp[r + 1]; p++) {
      int32_t q = fill[ci[p]]++;
      tci[q] = (int32_t)r;
      tv[q] = v[p];
    }
  Ctx *c = fm.ctx;
  F.nc = nc;
  F.P_rp.upload(c, rp);
  F.P_ci.upload(c, ci);
  F.R_rp.upload(c, trp);
  F.R_ci.upload(c, tci);
  F.P64.upload(c, v);
  F.R64.upload(c, tv);
  std::vector<float> v32(v.begin(), v.end()), tv32(tv.begin(), tv.end());
  F.P32.upload(c, v32);
  F.R32.upload(c, tv32);
}

// one fluid in a box (prescribed pressures allowed: the box solve then
// leaves their rows out, and the rediscretized hierarchy is not used)
static bool mg_possible(const Model &m) {
  if (!m.box.set || m.mats.size() != 1 || !m.mats.begin()->second.fluid) return false;
  for (auto v : m.row_field)
    if (v == 0) return false;
  return true;
}

// the 32-slot FP32 operator has variants (row kernel, row groups, tiles)
static bool variants_wanted(const elp_solve_opts &o) { return o.precision != ELP_PREC_FP64 && o.block == 32; }

// structure and fluid rows in one model
static bool two_fields(const Model &m) {
  bool has_s = false, has_f = false;
  for (auto v : m.row_field) {
    has_s |= v == 0;
    has_f |= v == 1;
  }
  return has_s && has_f;
}

// one fluid material and no structure rows: a mesh without a box of its own
// gets a stand-in box for the two-level cycle
static bool fluid_mesh(const Model &m) {
  if (m.mats.size() != 1 || !m.mats.begin()->second.fluid) return false;
  for (auto v : m.row_field)
    if (v == 0) return false;
  return true;
}

// The stand-in box of a fluid mesh (of a model's fluid): the bounding box of
// its fluid nodes with a uniform Hex8
// lattice of about the mesh's element size ((box volume / elements)^(1/3)),
// at most kFdmMaxAxis points across x and y, Robin faces on its z sides
// where the mesh has Robin faces lying on them (others are left to the
// smoothing steps and the Krylov iteration).
static BoxInfo standin_box(const Model &m) {
  BoxInfo b;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t r = 0; r < m.n_nodes; r++) {
    if (m.node_row[r] < 0 || m.row_field[m.node_row[r]] != 1) continue;  // fluid nodes
    const double *x = &m.xyz[(int64_t)m.rank_to_input[r] * 3];
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], x[d]);
      hi[d] = std::max(hi[d], x[d]);
    }
  }
  int64_t ne = 0;
  for (auto &kb : m.kinds)
    if (is_fluid(kb.kind)) ne += kb.n_elem;
  double vol = 1;
  for (int d = 0; d < 3; d++) vol *= hi[d] - lo[d];
  if (!(vol > 0) || ne <= 0) return b;
  const double h = std::cbrt(vol / (double)ne);
  for (int d = 0; d < 3; d++) {
    const int cap = d < 2 ? kFdmMaxAxis - 1 : 4096;
    b.origin[d] = lo[d];
    b.length[d] = hi[d] - lo[d];
    b.nel[d] = (int)std::min<double>(cap, std::max(1.0, std::round(b.length[d] / h)));
  }
  const double tol = 1e-9 * b.length[2];
  for (int64_t f = 0; f < m.n_faces; f++) {
    bool on_lo = true, on_hi = true;
    for (int k = 0; k < m.nfn; k++) {
      const double z = m.xyz[(int64_t)m.rank_to_input[m.face_nodes[f * m.nfn + k]] * 3 + 2];
      on_lo &= std::fabs(z - lo[2]) <= tol;
      on_hi &= std::fabs(z - hi[2]) <= tol;
    }
    if (on_lo) b.robin_sides |= 1 << 2;
    if (on_hi) b.robin_sides |= 1 << 4;
  }
  b.order = 2;  // the smoothing two-level cycle
  b.set = true;
  return b;
}

// structure rows beside one fluid: the coupled preconditioner (the fluid
// rows through the exact box solve of a Hex8 box given or read from the
// mesh, or the two-level cycle on a stand-in box)
static bool coupled_possible(const Model &m) {
  if (!two_fields(m)) return false;
  int n_fluid = 0;
  for (auto &kv : m.mats) n_fluid += kv.second.fluid;
  return n_fluid == 1;
}

static void choose_operator(Model &m, Model::SolverState &st, const elp_solve_opts &o, hipStream_t s);

static Model::SolverState &solver_state(Model &m, const elp_solve_opts &o, std::complex<double> Z, hipStream_t s) {
  const bool want_mg = o.precond == ELP_PC_MG && mg_possible(m);
  const bool want_cpl = o.precond == ELP_PC_MG && !want_mg && coupled_possible(m);
  if (m.solver && m.solver->mg == want_mg && m.solver->mg_levels == o.mg_levels &&
      m.solver->mg_min_nodes == o.mg_min_nodes && m.solver->cpl_requested == want_cpl &&
      m.solver->precond == o.precond)
    return *m.solver;
  m.solver.reset();
  auto st = std::make_shared<Model::SolverState>();
  st->mg = want_mg;
  st->precond = o.precond;
  st->cpl_requested = want_cpl;
  st->mg_levels = o.mg_levels;
  st->mg_min_nodes = o.mg_min_nodes;
  st->fp64_inner = two_fields(m);
  auto L0 = std::make_unique<LevelData>();
  L0->model = &m;
  prepare_level(*L0, s);
  st->levels.push_back(std::move(L0));
  st->box = m.box;
  if (want_mg) {
    st->fdm = fdm_build(m, st->box, 0);
  } else if (o.precond == ELP_PC_MG && fluid_mesh(m)) {
    // a fluid mesh that is no box: the two-level cycle on a stand-in box
    const BoxInfo b = standin_box(m);
    if (b.set) st->fdm = fdm_build(m, b, 0);
    if (st->fdm) {
      st->box = b;
      st->standin = true;
    }
  }
  st->order = st->box.order;
  if (want_cpl) {
    LevelData &L = *st->levels[0];
    if (m.box.set && m.box.order == 1) st->fdm = fdm_build(m, true);  // the cavity is the box
    if (!st->fdm) {
      const BoxInfo b = standin_box(m);
      if (b.set) st->fdm = fdm_build(m, b, 2);
      if (st->fdm) {
        st->box = b;
        st->standin = true;
      }
    }
    st->order = st->box.order;
    if (st->fdm) {
      Model lattice;
      lattice.box = st->box;
      build_interp(L, lattice, st->box.nel, true);
    }
    permute_level_float(L, s);
    std::vector<int32_t> ip;
    if (L.permuted) L.iperm.download(ip);
    if (st->fdm) {
      st->bcr = bcr_plan(m, L.permuted ? &ip : nullptr);
      st->coupled = st->bcr != nullptr;
    }
    if (!st->coupled) st->fdm.reset();  // l1-Jacobi
  } else if (st->fdm) {
    // prolongation from the Hex8 lattice of the box to the model's rows
    Model lattice;
    lattice.box = st->box;
    // rows of prescribed pressures take no share of the box correction
    build_interp(*st->levels[0], lattice, st->box.nel, !m.dir_rows.empty());
    permute_level_float(*st->levels[0], s);
    if (st->box.order == 2) build_box_products(*st->levels[0], s);
  } else if (!want_mg || !m.dir_rows.empty()) {
    // l1-Jacobi (the rediscretized hierarchy knows no prescribed values)
    permute_level_float(*st->levels[0], s);
  } else if (want_mg) {
    int nel[3] = {m.box.nel[0], m.box.nel[1], m.box.nel[2]};
    bool first = true;
    const int max_levels = o.mg_levels > 0 ? o.mg_levels : 8;
    while ((int)st->levels.size() < max_levels) {
      int nc[3] = {nel[0], nel[1], nel[2]};
      if (!(first && m.box.order == 2)) {
        // halve (rounding up: the trilinear interpolation does not need nested grids)
        for (int d = 0; d < 3; d++) nc[d] = std::max(1, (nel[d] + 1) / 2);
        if ((int64_t)(nc[0] + 1) * (nc[1] + 1) * (nc[2] + 1) < o.mg_min_nodes) break;
      }
      first = false;
      auto cm = make_box_level(m, nc, Z);
      finalize_model(*cm);
      assemble_model(*cm, s);
      LevelData &fine = *st->levels.back();
      cm->box = m.box;
      build_interp(fine, *cm, nc);
      auto Lc = std::make_unique<LevelData>();
      Lc->own = std::move(cm);
      Lc->model = Lc->own.get();
      prepare_level(*Lc, s);
      st->levels.push_back(std::move(Lc));
      for (int d = 0; d < 3; d++) nel[d] = nc[d];
    }
  }
  ELP_HIP(hipStreamSynchronize(s));
  m.solver = st;
  return *st;
}

// ---------------------------------------------------------------------------
// work buffers and the iteration, per precision T and block size B
// ---------------------------------------------------------------------------
template <typename T>
struct LevelWork {
  using V = typename VT<T>::t;
  DArr<V> x, b, r, e, ta, tb;
};

static OpCoef make_coef(const Model &m, double w, std::complex<double> Z, double shift);

static bool same_opts(const elp_solve_opts &a, const elp_solve_opts &b) {
  return a.precision == b.precision && a.method == b.method && a.precond == b.precond && a.block == b.block &&
         a.rtol == b.rtol && a.max_iters == b.max_iters && a.inner_rtol == b.inner_rtol && a.shift == b.shift &&
         a.smooth == b.smooth && a.coarse_sweeps == b.coarse_sweeps && a.mg_levels == b.mg_levels &&
         a.mg_min_nodes == b.mg_min_nodes && a.op_tile_rows == b.op_tile_rows && a.graph == b.graph;
}

// models up to this many rows replay their FP32 iterations from a captured
// HIP graph by default (their kernels are short next to launch overhead)
constexpr int64_t kGraphAutoRows = 100000;

// Launch of the 32-slot FP32 operator of a level: on its row tiles or row
// groups when it has them, else the row kernel (all give the same bits)
constexpr int kTgMaxLines = kTileMaxLines;  // 191 or 127 lines (more workgroups per CU) measured slower
template <int MODE>
static void launch_row_op(const LevelData &L, const OpCoef *coef, const float *x, const float *b, float *y,
                          hipStream_t s) {
  const OpArgs<float> a = L.template args<float>();
  if (L.n_tiles > 0 && L.t_half) {
    hipLaunchKernelGGL((thgrp_op_kernel<MODE>), dim3(xcd_grid(2 * L.n_tiles)), dim3(kTgWg),
                       (size_t)(L.t_lines + 1) * 128, s, a, L.padded(), (int64_t)L.pad_K.n, L.tiles(), L.t_grp.p,
                       L.g_rows.p, coef, x, b, y);
  } else if (L.n_tiles > 0 && (L.g_R == 2 || L.g_R == 4)) {
    const dim3 grid(xcd_grid(L.n_tiles));
    const size_t lds = (size_t)(L.t_lines + 1) * 256;
    if (L.g_R == 2)
      hipLaunchKernelGGL((tgrp_op_kernel<MODE, 2>), grid, dim3(kTgWg), lds, s, a, L.padded(), (int64_t)L.pad_K.n,
                         L.tiles(), L.t_grp.p, L.g_rows.p, coef, x, b, y);
    else
      hipLaunchKernelGGL((tgrp_op_kernel<MODE, 4>), grid, dim3(kTgWg), lds, s, a, L.padded(), (int64_t)L.pad_K.n,
                         L.tiles(), L.t_grp.p, L.g_rows.p, coef, x, b, y);
  } else if (L.n_tiles > 0) {
    hipLaunchKernelGGL((tile_op_kernel<MODE>), dim3(xcd_grid(L.n_tiles)), dim3(kTileWg),
                       (size_t)(L.t_lines + 1) * 256, s, a, L.padded(), L.tiles(), coef, x, b, y);
  } else if (L.g_R == 2 || L.g_R == 4) {
    const dim3 grid(xcd_grid(nblk(L.g_waves, kRowWg / 64)));
    if (L.g_R == 2)
      hipLaunchKernelGGL((rgrp_op_kernel<MODE, 2>), grid, dim3(kRowWg), 0, s, a, L.padded(), (int64_t)L.pad_K.n,
                         L.g_rows.p, L.g_waves, coef, x, b, y);
    else
      hipLaunchKernelGGL((rgrp_op_kernel<MODE, 4>), grid, dim3(kRowWg), 0, s, a, L.padded(), (int64_t)L.pad_K.n,
                         L.g_rows.p, L.g_waves, coef, x, b, y);
  } else {
    hipLaunchKernelGGL((row_op_kernel<MODE>), dim3(xcd_grid(nblk(L.n, kRowWg / 64))), dim3(kRowWg), 0, s, a,
                       L.padded(), coef, x, b, y);
  }
  ELP_HIP(hipGetLastError());
}

// the box cycle's fused post-smoothing: on the grouped half-line tiles when
// the set-up chose them, else one row per wavefront (the same bits)
static void launch_box_post(const LevelData &L, const OpCoef *coef, const float *e, int64_t ne, const float *u,
                            const float *v, float *z, hipStream_t s, bool row_kernels = false) {
  if (L.bt_on && !row_kernels) {
    hipLaunchKernelGGL(thbox_post_kernel, dim3(xcd_grid(2 * L.n_btiles)), dim3(kTgWg), (size_t)(L.bt_lines + 1) * 128,
                       s, L.template args<float>(), L.box_tiles(), coef, e, u, v, z);
  } else {
    hipLaunchKernelGGL(box_post_kernel, dim3(xcd_grid(nblk(L.n, kRowWg / 64))), dim3(kRowWg), 0, s,
                       L.template args<float>(), L.ap(), L.pp_off.p, L.pp_v.p, coef, e, ne, u, v, z);
  }
  ELP_HIP(hipGetLastError());
}

// The post-smoothing kernel of level 0 (Hex20 box cycle): the grouped
// half-line tiles (of 32, 64 or 128 rows) when they build and are faster on
// this model than one row per wavefront, each timed over 3 launches (after
// one untimed) on vectors of ones with the coefficients of a mid-range
// frequency.
static void choose_box_post(Model &m, Model::SolverState &st, hipStream_t s) {
  LevelData &L = *st.levels[0];
  if (st.box_chosen || !L.ap_rp.p || !st.fdm) return;
  st.box_chosen = true;
  if (!build_box_tiles(L, 32)) return;
  const int64_t ne = fdm_nodes(*st.fdm);
  DArr<float> ev, uv, vv, zv;
  ev.alloc(m.ctx, ne * 64);
  uv.alloc(m.ctx, m.n * 64);
  vv.alloc(m.ctx, m.n * 64);
  zv.alloc(m.ctx, m.n * 64);
  std::vector<float> ones((size_t)std::max<int64_t>(ne, m.n) * 64, 1.0f);
  ELP_HIP(hipMemcpyAsync(ev.p, ones.data(), sizeof(float) * ne * 64, hipMemcpyHostToDevice, s));
  ELP_HIP(hipMemcpyAsync(uv.p, ones.data(), sizeof(float) * m.n * 64, hipMemcpyHostToDevice, s));
  ELP_HIP(hipMemcpyAsync(vv.p, ones.data(), sizeof(float) * m.n * 64, hipMemcpyHostToDevice, s));
  std::vector<OpCoef> hc(32, make_coef(m, 2 * M_PI * 500.0, m.Z, 0.0));
  DArr<OpCoef> dc;
  dc.upload(m.ctx, hc);
  HipEvent e0, e1;
  auto time_it = [&]() {
    launch_box_post(L, dc.p, ev.p, ne, uv.p, vv.p, zv.p, s);
    ELP_HIP(hipEventRecord(e0.e, s));
    for (int k = 0; k < 3; k++) launch_box_post(L, dc.p, ev.p, ne, uv.p, vv.p, zv.p, s);
    ELP_HIP(hipEventRecord(e1.e, s));
    ELP_HIP(hipEventSynchronize(e1.e));
    float ms = 0;
    ELP_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
    return ms;
  };
  L.bt_on = false;
  float best = time_it();  // one row per wavefront
  int best_rows = 0, built = 32;
  for (int rows : {32, 64, 128}) {
    if (rows != 32) {
      if (!build_box_tiles(L, rows)) break;
      built = rows;
    }
    L.bt_on = true;
    const float ms = time_it();
    if (ms < best) {
      best = ms;
      best_rows = rows;
    }
  }
  if (best_rows == 0) release_box_tiles(L);
  else if (best_rows != built && !build_box_tiles(L, best_rows)) release_box_tiles(L);
  L.bt_on = best_rows != 0 && L.n_btiles > 0;
  L.bt_rows = best_rows;
}

// the variants of the 32-slot FP32 operator (elp_solve_opts.op_tile_rows):
// -1 the row kernel, -2 and -4 rows grouped 2 or 4 to a wavefront, t > 0 LDS
// tiles of t rows, 1000 R + t tiles of t rows with rows grouped R (2, 4)
// 6000 + t: tiles of t rows grouped 4 with each tile's frequencies split
// over two workgroups
static bool operator_variant_valid(int v) {
  if (v == 0 || v == -1 || v == -2 || v == -4) return true;
  if (v > 0 && v < 1000) return true;
  const int R = v / 1000, t = v % 1000;
  return (R == 2 || R == 4 || R == 6) && t > 0;
}
// the variant a level runs (the op_tile_rows encoding)
static int operator_variant(const LevelData &L) {
  if (L.n_tiles > 0 && L.t_half) return 6000 + L.t_rows;
  if (L.n_tiles > 0) return L.g_R > 1 ? 1000 * L.g_R + L.t_rows : L.t_rows;
  return L.g_R > 1 ? -L.g_R : -1;
}

// Set level 0 up for operator variant v: tiles and groups of another
// variant released, the padded copy rebuilt when its padding changes.
static void set_operator_variant(LevelData &L, int v, const std::vector<int32_t> &perm,
                                 const std::vector<int32_t> &iperm, const std::vector<int32_t> &rp, hipStream_t s) {
  L.n_tiles = 0;
  L.t_rows = 0;
  L.t_row.release();
  L.t_cptr.release();
  L.t_col.release();
  L.t_loff.release();
  L.t_grp.release();
  L.t_half = false;
  L.g_R = 0;
  L.g_waves = 0;
  L.g_rows.release();
  if (v >= 6000) {
    build_tiles(L, perm, iperm, rp, v % 1000, kTgMaxLines, 4, s, 128);
    L.t_half = L.n_tiles > 0;
    return;
  }
  if (v <= -2) {
    std::vector<int32_t> grows;
    const std::vector<int32_t> rp8 = group_padding(rp, -v, grows);
    pad_operator(L, rp8, -v, s);
    L.g_rows.upload(L.model->ctx, grows);
    L.g_waves = (int64_t)grows.size() / -v;
    L.g_R = -v;
    return;
  }
  if (v > 0) {
    build_tiles(L, perm, iperm, rp, v % 1000, v >= 1000 ? kTgMaxLines : kTileMaxLines, v / 1000, s);
    return;
  }
  if (L.pad_R != 0) pad_operator(L, row_padding(rp), 0, s);
}

// The 32-slot FP32 operator of level 0: the variant requested, or
// (op_tile_rows = 0) whichever of the row kernel, rows grouped 4 to a
// wavefront, row tiles of 32 rows, row tiles of 32 or 64 rows grouped 2 or 4
// to a wavefront and those grouped 4 with their frequencies split over two
// workgroups is fastest on this model:
// each candidate is timed over 3 launches (after one untimed) on a vector of
// ones with the coefficients of a mid-range frequency.
static void choose_operator(Model &m, Model::SolverState &st, const elp_solve_opts &o, hipStream_t s) {
  LevelData &L = *st.levels[0];
  if (st.fp64_inner || !variants_wanted(o) || !L.pad_rp.p) return;
  if (st.op_chosen && st.op_tile_rows == o.op_tile_rows) return;
  st.op_chosen = true;
  st.op_tile_rows = o.op_tile_rows;
  std::vector<int32_t> perm, iperm, rp;
  L.perm.download(perm);
  L.iperm.download(iperm);
  L.rp32.download(rp);
  if (o.op_tile_rows != 0) {
    set_operator_variant(L, o.op_tile_rows, perm, iperm, rp, s);
    return;
  }

  const int64_t nt = m.n * 64;
  DArr<float> xv, yv;
  xv.alloc(m.ctx, nt);
  yv.alloc(m.ctx, nt);
  std::vector<float> ones(nt, 1.0f);
  ELP_HIP(hipMemcpyAsync(xv.p, ones.data(), sizeof(float) * nt, hipMemcpyHostToDevice, s));
  std::vector<OpCoef> hc(32, make_coef(m, 2 * M_PI * 500.0, m.Z, 0.0));
  DArr<OpCoef> dc;
  dc.upload(m.ctx, hc);
  HipEvent e0, e1;
  auto time_it = [&]() {
    launch_row_op<0>(L, dc.p, xv.p, nullptr, yv.p, s);
    ELP_HIP(hipEventRecord(e0.e, s));
    for (int k = 0; k < 3; k++) launch_row_op<0>(L, dc.p, xv.p, nullptr, yv.p, s);
    ELP_HIP(hipEventRecord(e1.e, s));
    ELP_HIP(hipEventSynchronize(e1.e));
    float ms = 0;
    ELP_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
    return ms;
  };
  float best = 0;
  int best_v = 0, last = 0;
  for (int v : {-1, -4, 32, 2032, 4032, 4064, 6032, 6064}) {
    set_operator_variant(L, v, perm, iperm, rp, s);
    last = v;
    if (v > 0 && L.n_tiles == 0) continue;  // rows too long for tiles
    const float ms = time_it();
    if (best_v == 0 || ms < best) {
      best = ms;
      best_v = v;
    }
  }
  if (best_v != last) set_operator_variant(L, best_v, perm, iperm, rp, s);
}

template <typename T, int B>
struct Solver {
  using V = typename VT<T>::t;
  Model &m;
  Model::SolverState &st;
  elp_solve_opts o;
  hipStream_t s;
  Ctx *ctx;
  std::vector<LevelWork<T>> lw;
  DArr<OpCoef> coefA, coefP;
  DArr<double2> red_part, red_out, rho, rho_new, mu, alpha, beta;
  DArr<int> active;
  DArr<V> xi, ri, zi, pi, qi;    // inner Krylov vectors
  DArr<V> fb, fe, fs1, fs2;      // exact box solve: rhs, solution, scratch (lattice size)
  DArr<FdmCoef> coefF;
  // coupled models: the structure block's factorization in use (at
  // frequency fac_hz) for every slot, in FP64 also for the FP32 inner solve:
  // the stiffness of a thin plate has cond ~1e7-1e9, beyond an FP32
  // factorization
  std::unique_ptr<BcrFactor<double>, void (*)(BcrFactor<double> *)> fac{nullptr, nullptr};
  double fac_hz = -1;
  int refactors = 0;
  double refactor_ms = 0;
  // the next window's factorization, prepared on a thread and stream of its
  // own while the current window iterates
  std::unique_ptr<BcrFactor<double>, void (*)(BcrFactor<double> *)> next{nullptr, nullptr};
  double next_hz = -1;
  std::thread pre;
  std::exception_ptr pre_err;
  hipStream_t s2 = nullptr;
  PinnedArr<double2> h_red;     // per-frequency reductions, pinned host memory
  // statistics
  HipEvent ev0, ev1;
  double spmv_ms = 0;
  int64_t spmv_calls = 0;
  bool timing = true;
  const V *u_of = nullptr;  // level-0 x holds D^-1 u_of (first smoothing sweep)
  // one COCG iteration captured in a HIP graph (FP32 sweeps of launch-bound
  // models): replayed instead of launching its ~25 kernels one by one
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int64_t graph_replays = 0;
  bool use_graph = false;

  Solver(Model &mm, Model::SolverState &ss, const elp_solve_opts &oo, hipStream_t ss2)
      : m(mm), st(ss), o(oo), s(ss2), ctx(mm.ctx) {
    const int nl = (int)st.levels.size();
    lw.resize(nl);
    for (int l = 0; l < nl; l++) {
      const int64_t n = st.levels[l]->n * B;
      lw[l].x.alloc(ctx, n);
      lw[l].b.alloc(ctx, n);
      lw[l].r.alloc(ctx, n);
      lw[l].e.alloc(ctx, n);
      lw[l].ta.alloc(ctx, n);
      lw[l].tb.alloc(ctx, n);
    }
    const int64_t n0 = m.n * B;
    xi.alloc(ctx, n0); ri.alloc(ctx, n0); zi.alloc(ctx, n0); pi.alloc(ctx, n0); qi.alloc(ctx, n0);
    coefA.alloc(ctx, B);
    coefP.alloc(ctx, B);
    red_part.alloc(ctx, (size_t)kRedBlocks * B);
    red_out.alloc(ctx, B);
    rho.alloc(ctx, B); rho_new.alloc(ctx, B); mu.alloc(ctx, B); alpha.alloc(ctx, B); beta.alloc(ctx, B);
    active.alloc(ctx, B);
    if (st.fdm) {
      const int64_t nf = fdm_nodes(*st.fdm) * B;
      fb.alloc(ctx, nf); fe.alloc(ctx, nf); fs1.alloc(ctx, nf); fs2.alloc(ctx, nf);
      coefF.alloc(ctx, B);
    }
    h_red.alloc(B);
  }

  template <int MODE, typename TT = T>
  void op(int l, const OpCoef *coef, const typename VT<TT>::t *x, const typename VT<TT>::t *b,
          typename VT<TT>::t *y) {
    const LevelData &L = *st.levels[l];
    if constexpr (MODE == 3) {
      hipLaunchKernelGGL((jacobi0_kernel<TT>), dim3(nblk(L.n * B, 256)), dim3(256), 0, s, L.template args<TT>(), B,
                         coef, b, y);
      ELP_HIP(hipGetLastError());
      return;
    }
    if constexpr (std::is_same<TT, float>::value && B == 32) {
      if (L.pad_rp.p) {
        launch_row_op<MODE>(L, coef, reinterpret_cast<const float *>(x), reinterpret_cast<const float *>(b),
                            reinterpret_cast<float *>(y), s);
        return;
      }
    }
    const int grid = xcd_grid(nblk(L.n, OpShape<B>::RPB));
    hipLaunchKernelGGL((op_kernel<TT, B, MODE>), dim3(grid), dim3(256), 0, s, L.template args<TT>(), coef, x, b, y);
    ELP_HIP(hipGetLastError());
  }
  // r64 = b64 - A x in FP64 on the nw.nfz frozen columns, x: their [row][CB]
  // gather; coef holds CB entries
  template <int CB>
  void residual_fp64_narrow(const OpCoef *coef, const double2 *x, const NarrowOut &nw) {
    const LevelData &L = *st.levels[0];
    hipLaunchKernelGGL((narrow_residual_kernel<CB>), dim3(xcd_grid(nblk(L.n, OpShape<CB>::RPB))), dim3(256), 0, s,
                       L.template args<double>(), coef, x, nw);
    ELP_HIP(hipGetLastError());
  }
  template <bool ADD>
  void transfer(int64_t n, const int32_t *rp, const int32_t *ci, const T *v, const V *x, V *y) {
    hipLaunchKernelGGL((transfer_kernel<T, B, ADD>), dim3(xcd_grid(nblk(n, OpShape<B>::RPB))), dim3(256), 0, s, n, rp,
                       ci, v, x, y);
    ELP_HIP(hipGetLastError());
  }
  const T *Pv(const LevelData &L) const;
  const T *Rv(const LevelData &L) const;

  // per-frequency reduction into h_red (host, after sync)
  template <typename VV>
  void reduce(int64_t n, const VV *a, const VV *b, int mode, double2 *dev_out) {
    hipLaunchKernelGGL(dot_kernel<VV>, dim3(red_blocks(n)), dim3(256), 0, s, n, B, a, b, mode, red_part.p);
    final_reduce(dev_out, red_blocks(n));
  }
  // ||v||^2 and ||D^-1/2 v||^2 per frequency into dev_out (.x, .y): v in
  // the inner solve's row order, or (wreduce64) an FP64 vector in the
  // model's row order
  void wreduce(const V *v, double2 *dev_out) {
    hipLaunchKernelGGL((wnorm_kernel<T, V>), dim3(red_blocks(m.n)), dim3(256), 0, s, m.n, B, v,
                       st.levels[0]->template args<T>(), coefA.p, red_part.p);
    final_reduce(dev_out, red_blocks(m.n));
  }
  void wreduce64(const double2 *v, double2 *dev_out) {
    hipLaunchKernelGGL((wnorm_kernel<double, double2>), dim3(red_blocks(m.n)), dim3(256), 0, s, m.n, B, v,
                       st.levels[0]->template args<double>(), coefA.p, red_part.p);
    final_reduce(dev_out, red_blocks(m.n));
  }
  void final_reduce(double2 *dev_out, int nblocks) {
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kRedFinalThreads), 0, s, nblocks, B, red_part.p, dev_out);
    ELP_HIP(hipGetLastError());
  }
  // x += alpha p, r -= alpha q, red_out = |r|^2 per frequency (level 0);
  // with track_u, also the preconditioner's first sweep u = D^-1 r into the
  // level-0 work vector x (box_cycle then skips it)
  void update_norm(bool track_u = false) {
    hipLaunchKernelGGL((cocg_update_norm_kernel<T>), dim3(red_blocks(m.n)), dim3(256), 0, s, m.n, B, alpha.p, pi.p, qi.p,
                       xi.p, ri.p, red_part.p, st.levels[0]->template args<T>(), coefP.p, coefA.p,
                       track_u ? lw[0].x.p : nullptr);
    final_reduce(red_out.p, red_blocks(m.n));
    u_of = track_u ? ri.p : nullptr;
  }
  // the fused box cycle (32 frequencies, Hex20 box with products A P) is used
  bool fused_box() const {
    return std::is_same<T, float>::value && B == 32 && st.fdm && st.order == 2 && std::max(1, o.smooth) == 1 &&
           st.levels[0]->ap_rp.p;
  }
  void fetch(const double2 *dev) {
    ELP_HIP(hipMemcpyAsync(h_red.p, dev, sizeof(double2) * B, hipMemcpyDeviceToHost, s));
    ELP_HIP(hipStreamSynchronize(s));
  }

  // smoothing sweeps on level l from xin (nullptr = zero) into out
  void smooth(int l, const V *b, const V *xin, int sweeps, V *out) {
    LevelWork<T> &W = lw[l];
    const V *cur = xin;
    for (int k = 0; k < sweeps; k++) {
      V *dst = (k == sweeps - 1) ? out : ((cur == W.ta.p) ? W.tb.p : W.ta.p);
      if (cur == nullptr) op<3>(l, coefP.p, nullptr, b, dst);
      else op<2>(l, coefP.p, cur, b, dst);
      cur = dst;
    }
  }

  // z = V-cycle(b) on level l
  void vcycle(int l, const V *b, V *z) {
    const int nl = (int)st.levels.size();
    if (l == nl - 1) {
      smooth(l, b, nullptr, nl == 1 ? std::max(1, o.smooth) : o.coarse_sweeps, z);
      return;
    }
    LevelWork<T> &W = lw[l];
    const LevelData &L = *st.levels[l];
    const int nu = std::max(1, o.smooth);
    smooth(l, b, nullptr, nu, W.x.p);
    op<1>(l, coefP.p, W.x.p, b, W.r.p);
    LevelWork<T> &C = lw[l + 1];
    transfer<false>(L.nc, L.Rcsr<T>().rp, L.Rcsr<T>().ci, Rv(L), W.r.p, C.b.p);
    vcycle(l + 1, C.b.p, C.e.p);
    transfer<true>(L.n, L.Pcsr<T>().rp, L.Pcsr<T>().ci, Pv(L), C.e.p, W.x.p);
    smooth(l, b, W.x.p, nu, z);
  }

  // two-level cycle with the exact separable solve on the Hex8 lattice of the
  // box: Hex20 models smooth on their own level around it, Hex8 models use it
  // directly (it is their operator up to the Robin lumping)
  void box_cycle(const V *b, V *z) {
    const LevelData &L = *st.levels[0];
    LevelWork<T> &W = lw[0];
    const int64_t nf = fdm_nodes(*st.fdm);
    if (st.order == 1) {
      transfer<false>(nf, L.Rcsr<T>().rp, L.Rcsr<T>().ci, Rv(L), b, fb.p);
      fdm_solve(*st.fdm, B, coefF.p, fb.p, fe.p, fs1.p, fs2.p, s);
      transfer<false>(L.n, L.Pcsr<T>().rp, L.Pcsr<T>().ci, Pv(L), fe.p, z);
      return;
    }
    if (o.smooth == 0) {
      // additive two-level: z = D^-1 b + P A_box^-1 R b (no fine-level operator)
      op<3>(0, coefP.p, nullptr, b, z);
      transfer<false>(nf, L.Rcsr<T>().rp, L.Rcsr<T>().ci, Rv(L), b, fb.p);
      fdm_solve(*st.fdm, B, coefF.p, fb.p, fe.p, fs1.p, fs2.p, s);
      transfer<true>(L.n, L.Pcsr<T>().rp, L.Pcsr<T>().ci, Pv(L), fe.p, z);
      return;
    }
    const int nu = std::max(1, o.smooth);
    if constexpr (std::is_same<T, float>::value && B == 32) {
      if (nu == 1 && L.ap_rp.p) {
        // u = D^-1 r, v = r - A u, e = box solve of R v, then the fused
        // z = u + P e + D^-1 (v - (A P) e); u may come with r (update_norm)
        if (u_of != b) smooth(0, b, nullptr, 1, W.x.p);
        op<1>(0, coefP.p, W.x.p, b, W.r.p);
        transfer<false>(nf, L.Rcsr<T>().rp, L.Rcsr<T>().ci, Rv(L), W.r.p, fb.p);
        fdm_solve(*st.fdm, B, coefF.p, fb.p, fe.p, fs1.p, fs2.p, s);
        // op_tile_rows = -1: one row per wavefront here too
        launch_box_post(L, coefP.p, reinterpret_cast<const float *>(fe.p), nf, reinterpret_cast<const float *>(W.x.p),
                        reinterpret_cast<const float *>(W.r.p), reinterpret_cast<float *>(z), s,
                        o.op_tile_rows == -1);
        return;
      }
    }
    smooth(0, b, nullptr, nu, W.x.p);
    if (st.coupled) {
      // the fluid block alone: no share of the structure rows in A u (their
      // part of z is the structure solve's)
      hipLaunchKernelGGL((keep_field_kernel<V>), dim3(nblk(L.n * B, 256)), dim3(256), 0, s, L.n * B, B,
                         st.levels[0]->template args<T>().field, 1, W.x.p);
      ELP_HIP(hipGetLastError());
    }
    op<1>(0, coefP.p, W.x.p, b, W.r.p);
    transfer<false>(nf, L.Rcsr<T>().rp, L.Rcsr<T>().ci, Rv(L), W.r.p, fb.p);
    fdm_solve(*st.fdm, B, coefF.p, fb.p, fe.p, fs1.p, fs2.p, s);
    transfer<true>(L.n, L.Pcsr<T>().rp, L.Pcsr<T>().ci, Pv(L), fe.p, W.x.p);
    smooth(0, b, W.x.p, nu, z);
  }

  // factor the structure block at f_hz: A_ss = ks K + ms M
  void refactor(double f_hz, std::complex<double> ks, std::complex<double> ms) {
    join_prefetch();
    const auto t0 = std::chrono::steady_clock::now();
    if (!fac) fac = bcr_new_factor<double>();
    bcr_factor<double>(m, *st.bcr, ks, ms, *fac, s);
    fac_hz = f_hz;
    refactors++;
    refactor_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  // start factoring at f_hz in the background (one at a time)
  void prefetch(double f_hz, std::complex<double> ks, std::complex<double> ms) {
    join_prefetch();
    if (!next) next = bcr_new_factor<double>();
    if (!s2) ELP_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
    next_hz = f_hz;
    pre_err = nullptr;
    const int dev = ctx->device;
    pre = std::thread([this, ks, ms, dev]() {
      try {
        ELP_HIP(hipSetDevice(dev));
        bcr_factor<double>(m, *st.bcr, ks, ms, *next, s2);
      } catch (...) {
        pre_err = std::current_exception();
      }
    });
  }
  void join_prefetch() {
    if (pre.joinable()) pre.join();
  }
  // use the prefetched factorization if it was made at f_hz
  bool take_prefetch(double f_hz) {
    if (!pre.joinable() || std::fabs(next_hz - f_hz) > 1e-9 * std::max(1.0, f_hz)) return false;
    const auto t0 = std::chrono::steady_clock::now();
    pre.join();
    refactor_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (pre_err) std::rethrow_exception(pre_err);
    std::swap(fac, next);
    fac_hz = next_hz;
    refactors++;
    return true;
  }
  // ready for another sweep on stream ss (a solver kept from the last one)
  void rearm(hipStream_t ss) {
    join_prefetch();
    s = ss;
    spmv_ms = 0;
    spmv_calls = 0;
    graph_replays = 0;
    refactors = 0;
    refactor_ms = 0;
    fac_hz = -1;
    next_hz = -1;
    u_of = nullptr;
  }
  ~Solver() {
    join_prefetch();
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (s2) (void)hipStreamDestroy(s2);
  }

  void precond(const V *r, V *z) {
    if (st.coupled) {
      // block diagonal: the box solve on the fluid rows (0 elsewhere), then
      // the direct solve of the structure block on its rows; Dirichlet rows
      // stay 0 (their residual is 0)
      box_cycle(r, z);
      bcr_apply<double, V>(*st.bcr, *fac, B, r, z, s, std::is_same<T, double>::value);
      return;
    }
    if (st.fdm) box_cycle(r, z);
    else if (st.levels.size() == 1) smooth(0, r, nullptr, std::max(2, o.smooth), z);  // l1-Jacobi: 2 sweeps
    else vcycle(0, r, z);
  }

  // COCG on A (level 0, precision T) for the block: solves A d = rhs (rhs in
  // ri on entry) from d = 0 until both ||r_f|| <= tol ||rhs_f|| and
  // ||D^-1/2 r_f|| <= tol ||D^-1/2 rhs_f|| for all active f.  iters[f]
  // accumulates iterations while f is active.
  void cocg(const std::vector<int> &act0, double tol, int maxit, std::vector<int> &iters) {
    const int64_t nt = m.n * B;
    std::vector<int> act = act0;
    ELP_HIP(hipMemsetAsync(xi.p, 0, nt * sizeof(V), s));
    wreduce(ri.p, red_out.p);
    fetch(red_out.p);
    std::vector<double2> nr0(B);
    for (int f = 0; f < B; f++) nr0[f] = {std::sqrt(h_red[f].x), std::sqrt(h_red[f].y)};
    ELP_HIP(hipMemcpyAsync(active.p, act.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    precond(ri.p, zi.p);
    ELP_HIP(hipMemcpyAsync(pi.p, zi.p, nt * sizeof(V), hipMemcpyDeviceToDevice, s));
    reduce(m.n, ri.p, zi.p, 0, rho.p);
    const int tpb = 256;
    for (int it = 0; it < maxit; it++) {
      if (timing) ELP_HIP(hipEventRecord(ev0.e, s));
      op<0>(0, coefA.p, pi.p, nullptr, qi.p);
      if (timing) ELP_HIP(hipEventRecord(ev1.e, s));
      spmv_calls++;
      reduce(m.n, pi.p, qi.p, 0, mu.p);
      hipLaunchKernelGGL(cocg_alpha_kernel, dim3(1), dim3(64), 0, s, B, rho.p, mu.p, active.p, alpha.p);
      update_norm();
      fetch(red_out.p);
      if (timing) {
        float ms = 0;
        ELP_HIP(hipEventElapsedTime(&ms, ev0.e, ev1.e));
        spmv_ms += ms;
      }
      bool any = false, changed = false;
      for (int f = 0; f < B; f++) {
        if (!act[f]) continue;
        iters[f]++;
        const double nr = inner_rel(h_red[f], nr0[f]);
        if (!(nr > tol) || !std::isfinite(nr) || nr > 1e6) {
          act[f] = 0;
          changed = true;
        } else {
          any = true;
        }
      }
      if (!any) break;
      if (changed) ELP_HIP(hipMemcpyAsync(active.p, act.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
      precond(ri.p, zi.p);
      reduce(m.n, ri.p, zi.p, 0, rho_new.p);
      hipLaunchKernelGGL(cocg_beta_kernel, dim3(1), dim3(64), 0, s, B, rho_new.p, rho.p, active.p, beta.p);
      hipLaunchKernelGGL((cocg_dir_kernel<V>), dim3(nblk(nt, tpb)), dim3(tpb), 0, s, nt, B, beta.p, zi.p, pi.p);
      ELP_HIP(hipGetLastError());
    }
  }
};

template <typename T, int B>
const T *Solver<T, B>::Pv(const LevelData &L) const {
  if constexpr (std::is_same<T, float>::value) return L.P32.p;
  else return L.P64.p;
}
template <typename T, int B>
const T *Solver<T, B>::Rv(const LevelData &L) const {
  if constexpr (std::is_same<T, float>::value) return L.R32.p;
  else return L.R64.p;
}

static OpCoef make_coef(const Model &m, double w, std::complex<double> Z, double shift) {
  const std::complex<double> I(0, 1);
  std::complex<double> kf = 1.0 / (w * w), mf = -(1.0 - I * shift);
  std::complex<double> ks = 1.0 + I * m.eta * (m.visco_type == 2 ? w : 1.0);
  std::complex<double> ms = -w * w * (1.0 - I * shift);
  if (m.has_rayleigh) {
    ks += I * w * m.ray_a;
    ms += I * w * m.ray_b;
  }
  std::complex<double> g = (m.n_faces && std::abs(Z) > 0) ? I / (w * Z) : 0.0;
  OpCoef c;
  c.kf = {kf.real(), kf.imag()};
  c.mf = {mf.real(), mf.imag()};
  c.ks = {ks.real(), ks.imag()};
  c.ms = {ms.real(), ms.imag()};
  c.g = {g.real(), g.imag()};
  c.akf = std::abs(kf);
  c.amf = std::abs(mf);
  c.aks = std::abs(ks);
  c.ams = std::abs(ms);
  c.ag = std::abs(g);
  return c;
}

// Weight of the l1-Jacobi steps of the box two-level cycle: the smoother
// applies u = r / (theta d), d the l1 row bound |k| sum|K| + |m| sum|M| +
// |g| sum|C_G|.  theta = 1/2 is the largest step that amplifies no error
// component of a symmetric positive definite operator; on the Hex20 cavity it
// cut the sweep from 30.2 to 20.9 COCG iterations per frequency (0.45 needed
// 23.9, and 0.4 did not finish a 256-frequency sweep within minutes).
constexpr double kBoxSmoothWeight = 0.5;
static void weight_l1(OpCoef &c, double theta) {
  c.akf *= theta;
  c.amf *= theta;
  c.aks *= theta;
  c.ams *= theta;
  c.ag *= theta;
}

// Nonzero prescribed values: g (the values on the Dirichlet rows, 0
// elsewhere) and the free rows coupled to a Dirichlet column, whose
// right-hand side is lifted per frequency (lift_rhs_kernel).  Off when every
// prescribed value is 0: then b = F and x starts at 0, as without Dirichlet
// rows.
struct Lift {
  bool on = false;
  DArr<double2> g;
  DArr<int32_t> rows;
  int64_t nrows = 0;
};

static void build_lift(const Model &m, Lift &L) {
  bool any = false;
  for (auto &v : m.dir_vals) any |= v != std::complex<double>(0, 0);
  if (!any) return;
  std::vector<double2> g(m.n, double2{0, 0});
  for (size_t k = 0; k < m.dir_rows.size(); k++) g[m.dir_rows[k]] = {m.dir_vals[k].real(), m.dir_vals[k].imag()};
  std::vector<int32_t> rows;
  for (int64_t r = 0; r < m.n; r++) {
    if (m.row_field[r] == 2) continue;
    bool touch = false;
    for (int32_t p = m.rowptr[r]; p < m.rowptr[r + 1] && !touch; p++) touch = m.row_field[m.colind[p]] == 2;
    for (int32_t k = m.x_rowptr[r]; k < m.x_rowptr[r + 1] && !touch; k++) touch = m.row_field[m.x_col[k]] == 2;
    if (touch) rows.push_back((int32_t)r);
  }
  L.on = true;
  L.g.upload(m.ctx, g);
  L.rows.upload(m.ctx, rows);
  L.nrows = (int64_t)rows.size();
}

static void launch_lift(const Model &m, const Lift &L, int B, const int *fresh, const OpCoef *coefA,
                        const double2 *F, double2 *b64, hipStream_t s) {
  if (!L.on || L.nrows == 0) return;
  hipLaunchKernelGGL(lift_rhs_kernel, dim3(nblk(L.nrows * B, 256)), dim3(256), 0, s, L.nrows, L.rows.p, B, fresh,
                     m.d_rowptr.p, m.d_colind.p, m.d_row_field.p, m.d_K.p, m.d_M.p, m.d_x_rowptr.p, m.d_x_col.p,
                     m.d_x_type.p, m.d_x_val.p, L.g.p, coefA, F, b64);
  ELP_HIP(hipGetLastError());
}

// Structure-block factorization of the coupled preconditioner at f_hz:
// A_ss = ks K + ms M with the model's loss factor and Rayleigh terms.
static void plate_coefs(const Model &m, double f_hz, std::complex<double> &ks, std::complex<double> &ms) {
  const OpCoef c = make_coef(m, 2.0 * M_PI * f_hz, std::complex<double>(0, 0), 0.0);
  ks = {c.ks.x, c.ks.y};
  ms = {c.ms.x, c.ms.y};
}

// Automatic complex shift of the preconditioner's operator, P(w) = K/w^2 -
// (1 - i beta) M + ...: none where the box solve inverts the model's fluid
// operator itself (a Hex8 model on its own box, no prescribed pressures, or
// the fluid block of a coupled model); otherwise none while the mesh resolves
// the wave (k h < 0.5), growing to 0.3 on coarse meshes where the lattice and
// the model's discretization disagree near resonances; 0.5 for the
// rediscretized hierarchy.  A shift there would only damp an exact inverse:
// on a 24x24x29 coupled cavity swept to 2 kHz it raised the iterations per
// frequency from 27 to 239, on a 10k-DOF Hex8 cavity from 8 to 61.
static double auto_shift(const Model &m, const Model::SolverState &st, double w) {
  if (!st.fdm) return 0.5;
  if (st.order == 1 && !st.standin && (st.coupled || m.dir_rows.empty())) return 0.0;
  double hmax = 0;
  for (int d = 0; d < 3; d++) hmax = std::max(hmax, st.box.length[d] / st.box.nel[d]);
  const double kh = w / fdm_c(*st.fdm) * hmax;
  return std::min(0.3, std::max(0.0, 0.1 * (kh - 0.5)));
}

// One block of up to B frequencies: FP64 refinement around the inner solver.
template <typename T, int B>
static void solve_block(Solver<T, B> &S, const double *freqs, int nb,
                        std::complex<double> Z, double *X, double *relres, int32_t *iters, int &refinements) {
  Model &m = S.m;
  const elp_solve_opts &o = S.o;
  hipStream_t s = S.s;
  std::vector<OpCoef> cA(B), cP(B);
  std::vector<FdmCoef> cF(B);
  for (int f = 0; f < B; f++) {
    const double w = 2.0 * M_PI * freqs[std::min(f, nb - 1)];
    const double beta = o.shift < 0 ? auto_shift(m, S.st, w) : o.shift;
    cA[f] = make_coef(m, w, Z, 0.0);
    cP[f] = make_coef(m, w, Z, beta);
    if (S.st.fdm && S.st.order == 2) weight_l1(cP[f], kBoxSmoothWeight);
    if (S.st.fdm) {
      const double rho = fdm_rho(*S.st.fdm), c = fdm_c(*S.st.fdm);
      const std::complex<double> sM = (1.0 - std::complex<double>(0, beta)) / (c * c);
      const std::complex<double> tG = std::abs(Z) > 0 ? std::complex<double>(0, rho) / (w * Z) : 0.0;
      cF[f] = {1.0 / (w * w), {sM.real(), sM.imag()}, {tG.real(), tG.imag()}, rho};
    }
  }
  ELP_HIP(hipMemcpyAsync(S.coefA.p, cA.data(), sizeof(OpCoef) * B, hipMemcpyHostToDevice, s));
  ELP_HIP(hipMemcpyAsync(S.coefP.p, cP.data(), sizeof(OpCoef) * B, hipMemcpyHostToDevice, s));
  if (S.st.fdm) ELP_HIP(hipMemcpyAsync(S.coefF.p, cF.data(), sizeof(FdmCoef) * B, hipMemcpyHostToDevice, s));
  if (S.st.coupled) {
    // the structure block factored at the block's middle frequency (the
    // frequency itself for one-frequency blocks)
    std::vector<double> fs(freqs, freqs + nb);
    std::sort(fs.begin(), fs.end());
    std::complex<double> ks, ms;
    plate_coefs(m, fs[nb / 2], ks, ms);
    S.refactor(fs[nb / 2], ks, ms);
  }
  // FP64 outer vectors
  const int64_t nt = m.n * B;
  DArr<double2> F, b64, x64, r64;
  std::vector<double2> hF(m.n);
  for (int64_t r = 0; r < m.n; r++) hF[r] = {m.F[r].real(), m.F[r].imag()};
  for (size_t k = 0; k < m.dir_rows.size(); k++) hF[m.dir_rows[k]] = {m.dir_vals[k].real(), m.dir_vals[k].imag()};
  F.upload(m.ctx, hF);
  b64.alloc(m.ctx, nt);
  x64.alloc(m.ctx, nt);
  r64.alloc(m.ctx, nt);
  hipLaunchKernelGGL(bcast_rhs_kernel, dim3(nblk(nt, 256)), dim3(256), 0, s, m.n, B, F.p, b64.p);
  ELP_HIP(hipMemsetAsync(x64.p, 0, nt * sizeof(double2), s));
  // the FP64 operator lives in the FP64 solver when T = float; otherwise S itself
  auto op64 = [&](const double2 *x, const double2 *b, double2 *y) { S.template op<1, double>(0, S.coefA.p, x, b, y); };
  // residuals relative to ||F|| (= ||b|| without prescribed values), plain
  // and Jacobi-scaled (see SlotRun::Slot)
  S.wreduce64(b64.p, S.red_out.p);
  S.fetch(S.red_out.p);
  std::vector<double> nb2(B), nrefs(B);
  for (int f = 0; f < B; f++) {
    nb2[f] = std::sqrt(S.h_red[f].x);
    nrefs[f] = std::sqrt(S.h_red[f].y);
  }
  std::vector<double> nref(nb2), nrep(nb2);
  Lift lift;
  build_lift(m, lift);
  if (lift.on) {
    std::vector<int> all(B, 1);
    DArr<int> d_all;
    d_all.upload(m.ctx, all);
    launch_lift(m, lift, B, d_all.p, S.coefA.p, F.p, b64.p, s);
    DArr<double2> r0;
    r0.alloc(m.ctx, nt);
    hipLaunchKernelGGL(slot_fresh_kernel, dim3(nblk(nt, 256)), dim3(256), 0, s, m.n, B, d_all.p, b64.p, lift.g.p,
                       m.d_row_field.p, x64.p, r0.p);
    ELP_HIP(hipGetLastError());
    // stop against the free rows' lifted right-hand side, report against the
    // whole lifted right-hand side (see Slot::ref in solve_slots)
    S.wreduce64(r0.p, S.red_out.p);
    S.fetch(S.red_out.p);
    double gnorm2 = 0;
    for (auto &v : m.dir_vals) gnorm2 += std::norm(v);
    for (int f = 0; f < B; f++) {
      const double w0 = std::sqrt(S.h_red[f].x);
      if (w0 > 0) nref[f] = w0;
      nrep[f] = std::sqrt(w0 * w0 + gnorm2);
      nrefs[f] = std::sqrt(S.h_red[f].y);
    }
  }
  std::vector<int> it(B, 0);
  // rel = max(relp, rels): the plain and the Jacobi-scaled relative residual
  std::vector<double> rel(B, 1.0), relp(B, 1.0), rels(B, 1.0);
  auto rel_from_red = [&]() {
    for (int f = 0; f < B; f++) {
      relp[f] = nref[f] > 0 ? std::sqrt(S.h_red[f].x) / nref[f] : 0.0;
      rels[f] = nrefs[f] > 0 ? std::sqrt(S.h_red[f].y) / nrefs[f] : 0.0;
      rel[f] = std::max(relp[f], rels[f]);
    }
  };
  DArr<double> scl;
  scl.alloc(m.ctx, B);
  const int max_outer = 30;
  const LevelData &L0 = *S.st.levels[0];
  const bool use_perm = std::is_same<T, float>::value && L0.permuted;
  const int32_t *perm = use_perm ? L0.perm.p : nullptr;
  const int32_t *iperm = use_perm ? L0.iperm.p : nullptr;
  std::vector<double> prev(B, INFINITY), wprev(B, 0.0);
  std::vector<int> stalled(B, 0), lastact(B, 0);
  for (int outer = 0; outer <= max_outer; outer++) {
    op64(x64.p, b64.p, r64.p);
    S.wreduce64(r64.p, S.red_out.p);
    S.fetch(S.red_out.p);
    rel_from_red();
    // a refinement step that did not reduce the residual is taken back and
    // the frequency is not refined further
    std::vector<double> undo(B, 0.0);
    bool any_undo = false;
    for (int f = 0; f < B; f++) {
      if (!lastact[f]) continue;
      if (!(rel[f] < prev[f])) {
        undo[f] = -wprev[f];
        stalled[f] = 1;
        any_undo = true;
      } else if (rel[f] > 0.5 * prev[f]) {
        stalled[f] = 1;  // keep the small gain, stop refining
      }
    }
    if (any_undo) {
      ELP_HIP(hipMemcpyAsync(scl.p, undo.data(), sizeof(double) * B, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL((add_scaled_kernel<typename VT<T>::t>), dim3(nblk(nt, 256)), dim3(256), 0, s, nt, B,
                         scl.p, S.xi.p, x64.p, iperm);
      op64(x64.p, b64.p, r64.p);
      S.wreduce64(r64.p, S.red_out.p);
      S.fetch(S.red_out.p);
      rel_from_red();
    }
    std::vector<int> act(B, 0);
    std::vector<double> inv(B, 0.0), wgt(B, 0.0);
    bool any = false;
    for (int f = 0; f < B; f++) {
      const double nr = relp[f] * nref[f];  // ||r64||: the inner right-hand side is r64 / ||r64||
      if (f < nb && !stalled[f] && rel[f] > o.rtol && std::isfinite(rel[f]) && nr > 0) {
        act[f] = 1;
        any = true;
        inv[f] = 1.0 / nr;
        wgt[f] = nr;
      }
    }
    if (!any || outer == max_outer) break;
    int total = 0;
    for (int f = 0; f < nb; f++) total += it[f];
    if (total >= o.max_iters * nb) break;
    refinements++;
    ELP_HIP(hipMemcpyAsync(scl.p, inv.data(), sizeof(double) * B, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((scale_to_kernel<typename VT<T>::t>), dim3(nblk(nt, 256)), dim3(256), 0, s, nt, B, scl.p,
                       r64.p, S.ri.p, perm);
    double tol = o.inner_rtol;
    if (std::is_same<T, double>::value) {
      // FP64 inner solve: aim directly below the target
      double worst = 0;
      for (int f = 0; f < B; f++)
        if (act[f]) worst = std::max(worst, rel[f]);
      tol = std::max(0.3 * o.rtol / worst, 1e-15);
    }
    S.cocg(act, tol, o.max_iters, it);
    ELP_HIP(hipMemcpyAsync(scl.p, wgt.data(), sizeof(double) * B, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((add_scaled_kernel<typename VT<T>::t>), dim3(nblk(nt, 256)), dim3(256), 0, s, nt, B, scl.p,
                       S.xi.p, x64.p, iperm);
    ELP_HIP(hipGetLastError());
    ELP_HIP(hipStreamSynchronize(s));
    for (int f = 0; f < B; f++) {
      prev[f] = rel[f];
      wprev[f] = wgt[f];
      lastact[f] = act[f];
    }
  }
  for (int f = 0; f < nb; f++) {
    if (relres) relres[f] = std::max(nrep[f] > 0 ? relp[f] * nref[f] / nrep[f] : relp[f], rels[f]);
    if (iters) iters[f] = it[f];
  }
  if (X) {
    DArr<double2> out;
    out.alloc(m.ctx, m.n * nb);
    hipLaunchKernelGGL(unpack_kernel, dim3(nblk(m.n * nb, 256)), dim3(256), 0, s, m.n, B, nb, x64.p, out.p);
    ELP_HIP(hipGetLastError());
    ELP_HIP(hipMemcpyAsync(X, out.p, sizeof(double2) * m.n * nb, hipMemcpyDeviceToHost, s));
  }
  ELP_HIP(hipStreamSynchronize(s));
}

// Continuous batching of a sweep on B slots (mixed precision).  Each slot
// holds one frequency; all slots advance together through one fused COCG
// iteration (one multi-frequency SpMV and one preconditioner application
// for the whole block).  When a slot's inner solve reaches its tolerance
// (inner_rtol, or just past rtol when one FP32 solve can reach it) it is
// frozen; once a third of the slots are frozen (or none is iterating) the
// frozen slots take an FP64 refinement step together: x64 += w d, r64 = b -
// A x64 in FP64.  A slot whose FP64 residual reaches rtol (or stops
// improving, or runs out of iterations) is written out and takes the next
// frequency of the sweep; the others restart their inner solve on the new
// residual.  Every column is computed independently of the others, so a
// frequency's result does not depend on which frequencies share the block.
template <typename T, int B>
struct SlotRun {
  using V = typename VT<T>::t;
  static constexpr int tpb = 256;
  Solver<T, B> &S;
  Model &m;
  const elp_solve_opts &o;
  hipStream_t s;
  const int64_t nf;
  const double *freqs;
  const std::complex<double> Z;
  double *X, *relres;
  int32_t *iters;
  const int64_t nt;
  int refinements = 0;
  bool done = false;
  // right-hand side: F (elPaSo's, with the prescribed values), per column b
  // (lifted when values are prescribed), FP64 solution and residual
  DArr<double2> F, b64, x64, r64, col;
  double nbn = 0, gnorm2 = 0;
  Lift lift;
  const int32_t *perm = nullptr, *iperm = nullptr;
  // host mirrors (modified only while the stream is idle)
  std::vector<OpCoef> cA, cP;
  std::vector<FdmCoef> cF;
  std::vector<int> act, fresh;
  std::vector<double> scale, wgt;
  // narrow FP64 residual of the frozen columns (allocated at first use)
  std::vector<int> fmap;
  std::vector<OpCoef> cAn;
  DArr<int> d_fmap, d_fresh;
  DArr<OpCoef> d_cAn;
  DArr<double2> xn;
  DArr<double> d_scale, d_wgt;
  struct Slot {
    int64_t f = -1;   // frequency index, -1 = empty
    int state = 0;    // 0 empty, 1 iterating, 2 frozen (inner solve done)
    bool broken = false;
    double prev = 1;  // FP64 relative residual before the current inner solve
    double w = 0;     // ||r64|| the inner right-hand side was scaled by
    double tol = 0;   // relative tolerance of the current inner solve
    // The stopping test is ||r64|| <= rtol ref and the reported residual is
    // ||r64|| / rep.  Without prescribed values ref = rep = ||F||.  With
    // lifted values the system solved is A' x = b (b = F - A[:,D] g on the
    // free rows, g on the Dirichlet rows), x64 = g holds the Dirichlet rows
    // exactly, so r64 lives on the free rows: ref = ||b|| of the free rows,
    // rep = ||b||.  The values' identity rows (scale 1) and the free rows
    // (fluid rows divided by rho w^2, structure rows of order E h) can differ
    // by many orders, so neither part may set the scale of the other.
    double ref = 0, rep = 0;
    // The Jacobi-scaled residual ||D^-1/2 r64|| (D the l1 row bound of
    // A(w)) is held to rtol as well, relative to ||D^-1/2 b|| of the free
    // rows (refs): within one field the two measures agree, but fields can
    // differ in scale by 15 orders (a coupled model's fluid rows against its
    // structure rows), and the plain norm then sees one field only.  prev is
    // the larger of the two relative residuals (relp, rels); n0 = the plain
    // and scaled norms of the inner right-hand side r64 / w.
    double refs = 0, relp = 1, rels = 1;
    double2 n0 = {1, 0};
  };
  std::vector<Slot> sl;
  int64_t next = 0;
  std::vector<int> it_f;
  // slots take frequencies from the highest down: high frequencies need the
  // most iterations, so the batch drains on the cheap ones at the end
  std::vector<int64_t> order;
  // Coupled models: the structure block is factored at one frequency for a
  // window of frequencies: the sweep's range cut into as many equal windows
  // as it has slot-sized groups of frequencies, each factored at its centre.
  // Frequencies are taken in descending order, so the windows are used one
  // after the other and the next one is factored in the background.  A slot
  // takes a frequency of the next window only when the current window's
  // frequencies are done: switched in the middle of a window, the slots'
  // frequencies iterated under a factorization a window away (on a 152k-DOF
  // coupled model over 10-2000 Hz 44 iterations per frequency instead of 25,
  // at most 354 instead of 46).
  bool cpl = false;
  double ftop = 0, fbot = 0, wwidth = 0;
  int64_t nwin = 1, cur_win = -1, want_win = -1;
  bool need_refactor = false;
  int freeze_batch = 1;

  SlotRun(Solver<T, B> &S_, int64_t nf_, const double *freqs_, std::complex<double> Z_, double *X_, double *relres_,
          int32_t *iters_)
      : S(S_), m(S_.m), o(S_.o), s(S_.s), nf(nf_), freqs(freqs_), Z(Z_), X(X_), relres(relres_), iters(iters_),
        nt(S_.m.n * B), cA(B), cP(B), cF(B), act(B, 0), fresh(B, 0), scale(B, 0.0), wgt(B, 0.0), fmap(B, 0),
        cAn(B), sl(B), it_f(nf_, 0), order(nf_) {
    std::vector<double2> hF(m.n);
    for (int64_t r = 0; r < m.n; r++) hF[r] = {m.F[r].real(), m.F[r].imag()};
    for (size_t k = 0; k < m.dir_rows.size(); k++) hF[m.dir_rows[k]] = {m.dir_vals[k].real(), m.dir_vals[k].imag()};
    double nb2 = 0;
    for (const auto &v : hF) nb2 += v.x * v.x + v.y * v.y;
    nbn = std::sqrt(nb2);
    if (!(nbn > 0)) {  // b = 0: x = 0 for every frequency
      if (X) std::fill(X, X + 2 * nf * m.n, 0.0);
      for (int64_t k = 0; k < nf; k++) {
        if (relres) relres[k] = 0.0;
        if (iters) iters[k] = 0;
      }
      done = true;
      return;
    }
    F.upload(m.ctx, hF);
    b64.alloc(m.ctx, nt);
    x64.alloc(m.ctx, nt);
    r64.alloc(m.ctx, nt);
    if (X) col.alloc(m.ctx, m.n);
    build_lift(m, lift);
    for (auto &v : m.dir_vals) gnorm2 += std::norm(v);
    hipLaunchKernelGGL(bcast_rhs_kernel, dim3(nblk(nt, tpb)), dim3(tpb), 0, s, m.n, B, F.p, b64.p);
    ELP_HIP(hipMemsetAsync(x64.p, 0, nt * sizeof(double2), s));
    ELP_HIP(hipMemsetAsync(r64.p, 0, nt * sizeof(double2), s));
    ELP_HIP(hipMemsetAsync(S.ri.p, 0, nt * sizeof(V), s));
    ELP_HIP(hipMemsetAsync(S.xi.p, 0, nt * sizeof(V), s));
    ELP_HIP(hipMemsetAsync(S.pi.p, 0, nt * sizeof(V), s));
    const LevelData &L0 = *S.st.levels[0];
    const bool use_perm = std::is_same<T, float>::value && L0.permuted;
    perm = use_perm ? L0.perm.p : nullptr;
    iperm = use_perm ? L0.iperm.p : nullptr;
    // slots left empty keep valid (unused) coefficients
    for (int k = 0; k < B; k++) set_coefs(k, 0);
    d_fmap.alloc(m.ctx, B);
    d_cAn.alloc(m.ctx, B);
    d_fresh.alloc(m.ctx, B);
    d_scale.alloc(m.ctx, B);
    d_wgt.alloc(m.ctx, B);
    for (int64_t k = 0; k < nf; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return freqs[a] > freqs[b]; });
    cpl = S.st.coupled;
    ftop = freqs[order[0]];
    fbot = freqs[order[nf - 1]];
    nwin = cpl ? std::max<int64_t>(1, (nf + B - 1) / B) : 1;
    wwidth = nwin > 1 ? (ftop - fbot) / (double)nwin : 0.0;
    freeze_batch = std::max(1, B / 3);
  }

  // coefficients of frequency fi in slot k
  void set_coefs(int k, int64_t fi) {
    const double w = 2.0 * M_PI * freqs[fi];
    const double beta = o.shift < 0 ? auto_shift(m, S.st, w) : o.shift;
    cA[k] = make_coef(m, w, Z, 0.0);
    cP[k] = make_coef(m, w, Z, beta);
    if (S.st.fdm && S.st.order == 2) weight_l1(cP[k], kBoxSmoothWeight);
    if (S.st.fdm) {
      const double rho = fdm_rho(*S.st.fdm), c = fdm_c(*S.st.fdm);
      const std::complex<double> sM = (1.0 - std::complex<double>(0, beta)) / (c * c);
      const std::complex<double> tG = std::abs(Z) > 0 ? std::complex<double>(0, rho) / (w * Z) : 0.0;
      cF[k] = {1.0 / (w * w), {sM.real(), sM.imag()}, {tG.real(), tG.imag()}, rho};
    }
  }
  int64_t window_of(double f) const {
    if (!(wwidth > 0)) return 0;
    return std::min<int64_t>(nwin - 1, std::max<int64_t>(0, (int64_t)std::floor((ftop - f) / wwidth)));
  }
  double centre_of(int64_t w) const { return nwin > 1 ? ftop - wwidth * (w + 0.5) : 0.5 * (ftop + fbot); }
  void use_window(int64_t w) {
    std::complex<double> ks, ms;
    const double fc = centre_of(w);
    if (!S.take_prefetch(fc)) {
      plate_coefs(m, fc, ks, ms);
      S.refactor(fc, ks, ms);
    }
    cur_win = w;
    if (w + 1 < nwin) {
      const double fn = centre_of(w + 1);
      plate_coefs(m, fn, ks, ms);
      S.prefetch(fn, ks, ms);
    }
  }
  // inner tolerance for a slot at FP64 relative residual rel: inner_rtol, or
  // just past the target when that is within reach of one FP32 solve
  double inner_tol(double rel) const {
    // FP64 inner solves aim straight below the target
    if (std::is_same<T, double>::value) return std::max(0.3 * o.rtol / rel, 1e-15);
    const double finish_at = 0.5 * o.rtol / rel;
    return finish_at >= 1e-4 ? std::min(o.inner_rtol, finish_at) : o.inner_rtol;
  }
  // the reported residual: the plain one relative to ||b|| (F with the
  // prescribed values) or the scaled one, whichever is larger
  static double reported(const Slot &q, double rp, double rs) { return std::max(rp * q.ref / q.rep, rs); }
  // slot q restarts its inner solve on r64 with these relative residuals
  void new_rhs(Slot &q, double rp, double rs) const {
    q.relp = rp;
    q.rels = rs;
    q.prev = std::max(rp, rs);
    q.tol = inner_tol(q.prev);
    q.w = rp * q.ref;
    q.n0 = {1.0, q.w > 0 ? rs * q.refs / q.w : 0.0};
  }
  // assign the next frequency to slot k (stream idle)
  // a slot of the current window still iterating (not counting slot k)
  bool window_busy(int k) const {
    for (int j = 0; j < B; j++)
      if (j != k && sl[j].state != 0 && window_of(freqs[sl[j].f]) == cur_win) return true;
    return false;
  }
  bool assign(int k) {
    if (next >= nf || (cpl && cur_win >= 0 && window_of(freqs[order[next]]) != cur_win && window_busy(k))) {
      // coupled: a frequency of the next window waits until the current
      // window's frequencies are done, so that every frequency iterates
      // under its own window's structure factorization
      sl[k] = Slot();
      act[k] = 0;
      return false;
    }
    const int64_t fi = order[next++];
    set_coefs(k, fi);
    if (cpl && cur_win >= 0 && window_of(freqs[fi]) != cur_win) {
      // the slots move on to the next window
      want_win = need_refactor ? std::min(want_win, window_of(freqs[fi])) : window_of(freqs[fi]);
      need_refactor = true;
    }
    sl[k] = Slot();
    sl[k].f = fi;
    sl[k].state = 1;
    sl[k].prev = 1.0;
    sl[k].w = nbn;
    sl[k].ref = sl[k].rep = nbn;
    sl[k].tol = inner_tol(1.0);
    fresh[k] = 1;
    act[k] = 2;
    scale[k] = 1.0 / nbn;
    return true;
  }
  void upload_coefs() {
    ELP_HIP(hipMemcpyAsync(S.coefA.p, cA.data(), sizeof(OpCoef) * B, hipMemcpyHostToDevice, s));
    ELP_HIP(hipMemcpyAsync(S.coefP.p, cP.data(), sizeof(OpCoef) * B, hipMemcpyHostToDevice, s));
    if (S.st.fdm) ELP_HIP(hipMemcpyAsync(S.coefF.p, cF.data(), sizeof(FdmCoef) * B, hipMemcpyHostToDevice, s));
  }
  void finish(int k, double rel) {
    const int64_t fi = sl[k].f;
    if (relres) relres[fi] = rel;
    if (iters) iters[fi] = it_f[fi];
    if (X) {
      hipLaunchKernelGGL(unpack_col_kernel, dim3(nblk(m.n, tpb)), dim3(tpb), 0, s, m.n, B, k, x64.p, col.p);
      ELP_HIP(hipGetLastError());
      ELP_HIP(hipMemcpyAsync(X + 2 * fi * m.n, col.p, sizeof(double2) * m.n, hipMemcpyDeviceToHost, s));
    }
  }
  // start: fill the slots, x64 = g (0 without prescribed values), r64 = b -
  // A x64, inner rhs = r64 / ||r64|| (= b / ||b|| without prescribed values)
  void start_columns() {
    ELP_HIP(hipMemcpyAsync(d_fresh.p, fresh.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    launch_lift(m, lift, B, d_fresh.p, S.coefA.p, F.p, b64.p, s);
    hipLaunchKernelGGL(slot_fresh_kernel, dim3(nblk(nt, tpb)), dim3(tpb), 0, s, m.n, B, d_fresh.p, b64.p,
                       lift.on ? lift.g.p : nullptr, m.d_row_field.p, x64.p, r64.p);
    bool any_fresh = false;
    for (int k = 0; k < B; k++) any_fresh |= fresh[k] != 0;
    if (any_fresh) {
      // the norms of each fresh column's right-hand side: the scaled one,
      // and with lifted values the plain one (not ||F||) too
      S.wreduce64(r64.p, S.red_out.p);
      S.fetch(S.red_out.p);
      for (int k = 0; k < B; k++) {
        if (!fresh[k]) continue;
        Slot &q = sl[k];
        const double w0 = std::sqrt(S.h_red[k].x), ws0 = std::sqrt(S.h_red[k].y);
        if (lift.on) {
          q.w = w0;
          q.ref = w0 > 0 ? w0 : nbn;
          q.rep = std::sqrt(w0 * w0 + gnorm2);
          q.relp = w0 / q.ref;
          scale[k] = w0 > 0 ? 1.0 / w0 : 0.0;
        }
        q.refs = ws0;
        q.rels = ws0 > 0 ? 1.0 : 0.0;
        q.prev = std::max(q.relp, q.rels);
        q.n0 = {1.0, q.w > 0 ? ws0 / q.w : 0.0};
        q.tol = inner_tol(std::max(q.prev, 1e-300));
      }
    }
    ELP_HIP(hipMemcpyAsync(d_scale.p, scale.data(), sizeof(double) * B, hipMemcpyHostToDevice, s));
    ELP_HIP(hipMemcpyAsync(S.active.p, act.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    // restarted columns also get the preconditioner's first sweep (u) when
    // the update keeps it for the others
    hipLaunchKernelGGL((slot_restart_kernel<T>), dim3(nblk(nt, tpb)), dim3(tpb), 0, s, nt, B, S.active.p,
                       d_scale.p, r64.p, S.ri.p, S.xi.p, perm, S.st.levels[0]->template args<T>(), S.coefP.p,
                       S.fused_box() ? S.lw[0].x.p : nullptr);
    ELP_HIP(hipGetLastError());
    std::fill(fresh.begin(), fresh.end(), 0);
  }
  void start() {
    if (done) return;
    if (cpl) cur_win = 0;
    for (int k = 0; k < B; k++) assign(k);
    upload_coefs();
    if (cpl) use_window(0);
    start_columns();
  }
  // another COCG iteration is due (else the run is over)
  bool iterating() {
    if (done) return false;
    int n_iter = 0;
    for (int k = 0; k < B; k++) n_iter += sl[k].state == 1;
    if (n_iter == 0) {
      ELP_HIP(hipStreamSynchronize(s));
      done = true;
    }
    return !done;
  }
  // after p = z + beta p: columns that (re)started iterate from now on (the
  // device flags by clear_restart_kernel, unless a captured iteration does
  // it: device = false)
  void restart_flags(bool device = true) {
    bool restarted = false;
    for (int k = 0; k < B; k++)
      if (act[k] == 2) {
        act[k] = 1;
        restarted = true;
      }
    if (restarted && device) {
      hipLaunchKernelGGL(clear_restart_kernel, dim3(1), dim3(64), 0, s, B, S.active.p);
      ELP_HIP(hipGetLastError());
    }
  }
  // host side of an iteration, with ||r_f||^2 of every column in S.h_red:
  // frozen columns take an FP64 refinement step together, finished
  // frequencies are written out and their slots take the next ones
  void after_norms();
};

template <typename T, int B>
void SlotRun<T, B>::after_norms() {
  int n_frozen = 0, n_iter = 0;
  bool changed = false;
  for (int k = 0; k < B; k++) {
    Slot &q = sl[k];
    if (q.state == 1) {
      it_f[q.f]++;
      const double nr = inner_rel(S.h_red[k], q.n0);
      const bool bad = !std::isfinite(nr) || nr > 1e6;
      if (bad || !(nr > q.tol) || it_f[q.f] >= o.max_iters) {
        q.state = 2;
        q.broken = bad;
        act[k] = 0;
        changed = true;
      }
    }
    n_frozen += q.state == 2;
    n_iter += q.state == 1;
  }
  if (changed) ELP_HIP(hipMemcpyAsync(S.active.p, act.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
  if (n_frozen == 0 || (n_frozen < freeze_batch && n_iter > 0)) return;

  // FP64 refinement of the frozen columns
  refinements++;
  std::vector<int> was_iter(B, 0);
  for (int k = 0; k < B; k++) was_iter[k] = sl[k].state == 1;
  for (int k = 0; k < B; k++) wgt[k] = (sl[k].state == 2 && !sl[k].broken) ? sl[k].w : 0.0;
  ELP_HIP(hipMemcpyAsync(d_wgt.p, wgt.data(), sizeof(double) * B, hipMemcpyHostToDevice, s));
  // x64 += w x_in on the frozen columns, then r64 = b - A x64 on them: on a
  // narrow block of 8 or 16 columns when few are frozen (the gathers of x
  // shrink with the block; the update and the gather into the block are one
  // pass)
  int nfz = 0;
  for (int k = 0; k < B; k++)
    if (sl[k].state == 2) fmap[nfz++] = k;
  if (B == 32 && nfz <= 16) {
    const int CB = nfz <= 8 ? 8 : 16;
    for (int f = nfz; f < CB; f++) fmap[f] = fmap[0];  // padding columns take a frozen one's coefficients
    for (int f = 0; f < CB; f++) cAn[f] = cA[fmap[f]];
    if (!xn.p) xn.alloc(m.ctx, m.n * 16);
    ELP_HIP(hipMemcpyAsync(d_fmap.p, fmap.data(), sizeof(int) * CB, hipMemcpyHostToDevice, s));
    ELP_HIP(hipMemcpyAsync(d_cAn.p, cAn.data(), sizeof(OpCoef) * CB, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((add_gather_kernel<V>), dim3(nblk(m.n * CB, tpb)), dim3(tpb), 0, s, m.n, B, CB, nfz,
                       d_fmap.p, d_wgt.p, S.xi.p, iperm, x64.p, xn.p);
    const NarrowOut nw = {d_fmap.p, nfz, B, b64.p, r64.p};
    if (CB == 8) S.template residual_fp64_narrow<8>(d_cAn.p, xn.p, nw);
    else S.template residual_fp64_narrow<16>(d_cAn.p, xn.p, nw);
  } else {
    hipLaunchKernelGGL((add_scaled_kernel<V>), dim3(nblk(nt, tpb)), dim3(tpb), 0, s, nt, B, d_wgt.p, S.xi.p, x64.p,
                       iperm);
    S.template op<1, double>(0, S.coefA.p, x64.p, b64.p, r64.p);
  }
  S.wreduce64(r64.p, S.red_out.p);
  S.fetch(S.red_out.p);
  std::vector<double> rel(B), relp(B), rels(B);
  for (int k = 0; k < B; k++) {
    relp[k] = std::sqrt(S.h_red[k].x) / (sl[k].ref > 0 ? sl[k].ref : nbn);
    rels[k] = sl[k].refs > 0 ? std::sqrt(S.h_red[k].y) / sl[k].refs : 0.0;
    rel[k] = std::max(relp[k], rels[k]);
  }
  // a step that did not reduce the residual is taken back
  bool any_undo = false;
  std::vector<int> undo(B, 0);
  for (int k = 0; k < B; k++) {
    wgt[k] = 0.0;
    if (sl[k].state != 2 || sl[k].broken) continue;
    if (!(rel[k] < sl[k].prev)) {
      undo[k] = 1;
      wgt[k] = -sl[k].w;
      any_undo = true;
    }
  }
  if (any_undo) {
    ELP_HIP(hipMemcpyAsync(d_wgt.p, wgt.data(), sizeof(double) * B, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((add_scaled_kernel<V>), dim3(nblk(nt, tpb)), dim3(tpb), 0, s, nt, B, d_wgt.p, S.xi.p, x64.p,
                       iperm);
    ELP_HIP(hipGetLastError());
  }
  bool any_new = false;
  for (int k = 0; k < B; k++) {
    Slot &q = sl[k];
    if (q.state != 2) continue;
    const bool kept_back = q.broken || undo[k];
    const bool fin = kept_back || rel[k] <= o.rtol || rel[k] > 0.5 * q.prev || it_f[q.f] >= o.max_iters;
    if (fin) {
      finish(k, kept_back ? reported(q, q.relp, q.rels) : reported(q, relp[k], rels[k]));
      any_new |= assign(k);
    } else {
      new_rhs(q, relp[k], rels[k]);
      q.state = 1;
      act[k] = 2;
      scale[k] = 1.0 / q.w;
    }
  }
  if (cpl)
    for (int k = 0; k < B; k++)
      if (sl[k].state == 0) any_new |= assign(k);
  if (need_refactor) {
    // slots in the middle of an inner solve keep their progress (x64 +=
    // w d) and restart on the new residual under the new factorization
    bool anyf = false;
    for (int k = 0; k < B; k++) {
      wgt[k] = was_iter[k] ? sl[k].w : 0.0;
      anyf |= was_iter[k] != 0;
    }
    if (anyf) {
      ELP_HIP(hipMemcpyAsync(d_wgt.p, wgt.data(), sizeof(double) * B, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL((add_scaled_kernel<V>), dim3(nblk(nt, tpb)), dim3(tpb), 0, s, nt, B, d_wgt.p, S.xi.p,
                         x64.p, iperm);
      ELP_HIP(hipGetLastError());
      S.template op<1, double>(0, S.coefA.p, x64.p, b64.p, r64.p);
      S.wreduce64(r64.p, S.red_out.p);
      S.fetch(S.red_out.p);
      for (int k = 0; k < B; k++) {
        if (!was_iter[k]) continue;
        Slot &q = sl[k];
        new_rhs(q, std::sqrt(S.h_red[k].x) / q.ref, q.refs > 0 ? std::sqrt(S.h_red[k].y) / q.refs : 0.0);
        act[k] = 2;
        scale[k] = q.w > 0 ? 1.0 / q.w : 0.0;
      }
    }
    use_window(want_win);
    need_refactor = false;
  }
  // the stream is idle here only after a sync: the host mirrors are reused
  ELP_HIP(hipStreamSynchronize(s));
  if (any_new) upload_coefs();
  start_columns();
}

// one COCG iteration of a run on its own: z = M r, rho' = r.z, beta (0 for
// restarting columns), p = z + beta p; q = A p, alpha = rho / p.q, x +=
// alpha p, r -= alpha q, ||r||^2 per column to the host
// the iteration's launches, in the order of run_iteration, with the restart
// flags cleared on the device; graph capture of them ends with the copy of
// the norms to the host (no synchronization inside a capture)
template <typename T, int B>
static void iteration_body(SlotRun<T, B> &R, bool events) {
  using V = typename VT<T>::t;
  Solver<T, B> &S = R.S;
  hipStream_t s = S.s;
  const int64_t n = S.m.n, nt = n * B;
  S.precond(S.ri.p, S.zi.p);
  S.reduce(n, S.ri.p, S.zi.p, 0, S.rho_new.p);
  hipLaunchKernelGGL(cocg_beta_kernel, dim3(1), dim3(64), 0, s, B, S.rho_new.p, S.rho.p, S.active.p, S.beta.p);
  hipLaunchKernelGGL((cocg_dir_kernel<V>), dim3(nblk(nt, 256)), dim3(256), 0, s, nt, B, S.beta.p, S.zi.p, S.pi.p);
  hipLaunchKernelGGL(clear_restart_kernel, dim3(1), dim3(64), 0, s, B, S.active.p);
  ELP_HIP(hipGetLastError());
  if (events) ELP_HIP(hipEventRecord(S.ev0.e, s));
  S.template op<0>(0, S.coefA.p, S.pi.p, nullptr, S.qi.p);
  if (events) ELP_HIP(hipEventRecord(S.ev1.e, s));
  S.reduce(n, S.pi.p, S.qi.p, 0, S.mu.p);
  hipLaunchKernelGGL(cocg_alpha_kernel, dim3(1), dim3(64), 0, s, B, S.rho.p, S.mu.p, S.active.p, S.alpha.p);
  S.update_norm(S.fused_box());
  ELP_HIP(hipMemcpyAsync(S.h_red.p, S.red_out.p, sizeof(double2) * B, hipMemcpyDeviceToHost, s));
}

// One COCG iteration from the captured graph: the first iteration in the
// steady state (the preconditioner's first sweep kept by the update when the
// fused box cycle runs) is captured, later ones replay it.  The same kernels
// on the same buffers as run_iteration: the same bits.
template <typename T, int B>
static bool run_iteration_graph(SlotRun<T, B> &R) {
  Solver<T, B> &S = R.S;
  if (S.fused_box() && S.u_of != S.ri.p) return false;  // not yet steady: run it eagerly
  hipStream_t s = S.s;
  if (!S.gexec) {
    ELP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      iteration_body(R, false);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    ELP_HIP(hipStreamEndCapture(s, &S.graph));
    ELP_HIP(hipGraphInstantiate(&S.gexec, S.graph, nullptr, nullptr, 0));
  }
  ELP_HIP(hipGraphLaunch(S.gexec, s));
  ELP_HIP(hipStreamSynchronize(s));
  R.restart_flags(false);  // the graph cleared them on the device
  S.u_of = S.fused_box() ? S.ri.p : nullptr;
  S.spmv_calls++;
  S.graph_replays++;
  return true;
}

template <typename T, int B>
static void run_iteration(SlotRun<T, B> &R) {
  Solver<T, B> &S = R.S;
  if constexpr (std::is_same<T, float>::value) {
    if (S.use_graph && run_iteration_graph(R)) return;
  }
  iteration_body(R, S.timing);
  ELP_HIP(hipStreamSynchronize(S.s));
  R.restart_flags(false);  // cleared on the device by the iteration
  S.spmv_calls++;
  if (S.timing) {
    float ms = 0;
    ELP_HIP(hipEventElapsedTime(&ms, S.ev0.e, S.ev1.e));
    S.spmv_ms += ms;
  }
}

// Continuous batching of one sweep on B slots (see SlotRun).
template <typename T, int B>
static void solve_slots(Solver<T, B> &S, int64_t nf, const double *freqs, std::complex<double> Z, double *X,
                        double *relres, int32_t *iters, int &refinements) {
  SlotRun<T, B> R(S, nf, freqs, Z, X, relres, iters);
  R.start();
  while (R.iterating()) {
    run_iteration(R);
    R.after_norms();
  }
  refinements = R.refinements;
}

// ---------------------------------------------------------------------------
// Grouped execution: the sweeps of several requests advance together, one
// COCG iteration of every request per step, on one stream.  The operator,
// the dot products, the direction and the update of all requests are single
// launches over a device table of per-request segments; each segment is
// computed with exactly the partition (rows per workgroup, reduction blocks)
// of the request's own launches, so a request's results are the same bits
// alone, in any mix and through elp_sweep.  Preconditioners and FP64
// refinement steps stay per request.
// ---------------------------------------------------------------------------
template <typename T>
struct RunSeg {
  int64_t n = 0;
  OpArgs<T> a{};
  PadCsr pc{};  // row-padded FP32 operator (B = 32) or empty
  const OpCoef *coefA = nullptr, *coefP = nullptr;
  typename VT<T>::t *ri = nullptr, *zi = nullptr, *pi = nullptr, *qi = nullptr, *xi = nullptr, *u = nullptr;
  double2 *red_part = nullptr, *red_out = nullptr, *rho = nullptr, *rho_new = nullptr, *mu = nullptr,
          *alpha = nullptr, *beta = nullptr;
  const int *active = nullptr;
  int op_blk0 = 0, op_nblk = 0;  // operator workgroups (multiples of 8)
  int e_blk0 = 0;                // element-wise workgroups of 256 threads
  int r_blk0 = 0, r_nblk = 0;    // reduction workgroups (red_blocks(n))
};

// segment of block b: the last segment whose first block is <= b
template <typename T>
__device__ inline int seg_of(const RunSeg<T> *__restrict__ segs, int G, int b, int which) {
  int g = 0;
  for (int k = 1; k < G; k++)
    if ((which == 0 ? segs[k].op_blk0 : which == 1 ? segs[k].e_blk0 : segs[k].r_blk0) <= b) g = k;
  return g;
}

template <typename T, int B, int WHICH>
__global__ __launch_bounds__(256) void grp_dot_kernel(const RunSeg<T> *__restrict__ segs, int G) {
  const RunSeg<T> &q = segs[seg_of(segs, G, blockIdx.x, 2)];
  const int blk = (int)blockIdx.x - q.r_blk0;
  if (WHICH == 0) dot_body(q.n, B, q.ri, q.zi, 0, q.red_part, blk, q.r_nblk);
  else dot_body(q.n, B, q.pi, q.qi, 0, q.red_part, blk, q.r_nblk);
}

template <typename T, int B, int WHICH>
__global__ __launch_bounds__(kRedFinalThreads) void grp_final_kernel(const RunSeg<T> *__restrict__ segs) {
  const RunSeg<T> &q = segs[blockIdx.x];
  dot_final_body(q.r_nblk, B, q.red_part, WHICH == 0 ? q.rho_new : WHICH == 1 ? q.mu : q.red_out);
}

template <typename T, int B>
__global__ void grp_beta_kernel(const RunSeg<T> *__restrict__ segs) {
  const RunSeg<T> &q = segs[blockIdx.x];
  beta_body(B, q.rho_new, q.rho, q.active, q.beta);
}

template <typename T, int B>
__global__ void grp_alpha_kernel(const RunSeg<T> *__restrict__ segs) {
  const RunSeg<T> &q = segs[blockIdx.x];
  alpha_body(B, q.rho, q.mu, q.active, q.alpha);
}

template <typename T, int B>
__global__ __launch_bounds__(256) void grp_dir_kernel(const RunSeg<T> *__restrict__ segs, int G) {
  const RunSeg<T> &q = segs[seg_of(segs, G, blockIdx.x, 1)];
  const int64_t i = (int64_t)(blockIdx.x - q.e_blk0) * blockDim.x + threadIdx.x;
  dir_body(q.n * B, B, q.beta, q.zi, q.pi, i);
}

template <typename T, int B>
__global__ __launch_bounds__(256) void grp_update_kernel(const RunSeg<T> *__restrict__ segs, int G) {
  const RunSeg<T> &q = segs[seg_of(segs, G, blockIdx.x, 2)];
  update_norm_body<T>(q.n, B, q.alpha, q.pi, q.qi, q.xi, q.ri, q.red_part, q.a, q.coefP, q.coefA, q.u,
                      (int)blockIdx.x - q.r_blk0, q.r_nblk);
}

// q = A p of every segment: the row-per-wavefront FP32 operator where the
// segment has its row-padded copy (32 slots), the chunked operator otherwise
template <typename T, int B>
__global__ __launch_bounds__(256) void grp_op_kernel(const RunSeg<T> *__restrict__ segs, int G) {
  const RunSeg<T> &q = segs[seg_of(segs, G, blockIdx.x, 0)];
  const int local = xcd_map((int)blockIdx.x - q.op_blk0, q.op_nblk);
  if constexpr (std::is_same<T, float>::value && B == 32) {
    if (q.pc.rp) {
      const int64_t r = __builtin_amdgcn_readfirstlane(local * (kRowWg / 64) + (int)(threadIdx.x >> 6));
      if (r >= q.n) return;
      row_apply<0>(q.a, q.pc, q.coefA, reinterpret_cast<const float *>(q.pi), nullptr,
                   reinterpret_cast<float *>(q.qi), r, threadIdx.x & 63);
      return;
    }
  }
  op_body<T, B, 0>(q.a, q.coefA, q.pi, nullptr, q.qi, local);
}

template <typename T, int B>
struct GroupRun : GroupItem {
  Model &m;
  std::vector<double> freqs;
  std::complex<double> Z;
  elp_solve_opts o;
  double *X, *relres;
  int32_t *iters;
  std::unique_ptr<Solver<T, B>> S;
  std::unique_ptr<SlotRun<T, B>> R;
  // the request's own stream (preconditioner, refinement), joined to the
  // group's stream through `ready` before each fused launch
  hipStream_t own = nullptr;
  std::unique_ptr<HipEvent> ready;
  GroupRun(Model &mm, const double *f, int64_t nf, std::complex<double> z, const elp_solve_opts &oo, double *x,
           double *rel, int32_t *it)
      : m(mm), freqs(f, f + nf), Z(z), o(oo), X(x), relres(rel), iters(it) {}
  // finish: the returned results and the solver's memory released
  ~GroupRun() override {
    if (own) (void)hipStreamSynchronize(own);
    R.reset();
    S.reset();
    ready.reset();
    if (own) (void)hipStreamDestroy(own);
  }
};

template <typename T, int B>
struct GroupDriverT : GroupDriver {
  Ctx *ctx;
  hipStream_t s;
  std::vector<std::unique_ptr<GroupRun<T, B>>> runs;
  DArr<RunSeg<T>> d_segs;
  std::vector<RunSeg<T>> h_segs;
  bool dirty = true;
  int op_blocks = 0, e_blocks = 0, r_blocks = 0;
  HipEvent ev0, ev1;
  GroupDriverT(Ctx *c, hipStream_t st) : ctx(c), s(st) {}

  void add(std::unique_ptr<GroupItem> it) override {
    std::unique_ptr<GroupRun<T, B>> g(static_cast<GroupRun<T, B> *>(it.release()));
    ELP_HIP(hipStreamCreateWithFlags(&g->own, hipStreamNonBlocking));
    g->ready.reset(new HipEvent);
    g->S->s = g->own;
    g->R.reset(new SlotRun<T, B>(*g->S, (int64_t)g->freqs.size(), g->freqs.data(), g->Z, g->X, g->relres,
                                 g->iters));
    g->R->start();
    runs.push_back(std::move(g));
    dirty = true;
  }
  bool empty() const override { return runs.empty(); }

  void rebuild() {
    h_segs.assign(runs.size(), RunSeg<T>());
    int ob = 0, eb = 0, rb = 0;
    for (size_t k = 0; k < runs.size(); k++) {
      Solver<T, B> &S = *runs[k]->S;
      const LevelData &L = *S.st.levels[0];
      RunSeg<T> &q = h_segs[k];
      q.n = S.m.n;
      q.a = L.template args<T>();
      if constexpr (std::is_same<T, float>::value && B == 32) {
        q.pc = L.pad_rp.p ? L.padded() : PadCsr{};
      }
      q.coefA = S.coefA.p;
      q.coefP = S.coefP.p;
      q.ri = S.ri.p;
      q.zi = S.zi.p;
      q.pi = S.pi.p;
      q.qi = S.qi.p;
      q.xi = S.xi.p;
      q.u = S.fused_box() ? S.lw[0].x.p : nullptr;
      q.red_part = S.red_part.p;
      q.red_out = S.red_out.p;
      q.rho = S.rho.p;
      q.rho_new = S.rho_new.p;
      q.mu = S.mu.p;
      q.alpha = S.alpha.p;
      q.beta = S.beta.p;
      q.active = S.active.p;
      const int rows_per_blk = (q.pc.rp ? kRowWg / 64 : OpShape<B>::RPB);
      q.op_blk0 = ob;
      q.op_nblk = xcd_grid(nblk(q.n, rows_per_blk));
      ob += q.op_nblk;
      q.e_blk0 = eb;
      eb += nblk(q.n * B, 256);
      q.r_blk0 = rb;
      q.r_nblk = red_blocks(q.n);
      rb += q.r_nblk;
    }
    op_blocks = ob;
    e_blocks = eb;
    r_blocks = rb;
    if ((int64_t)d_segs.n < (int64_t)h_segs.size()) d_segs.alloc(ctx, std::max<size_t>(h_segs.size(), 16));
    ELP_HIP(hipMemcpyAsync(d_segs.p, h_segs.data(), sizeof(RunSeg<T>) * h_segs.size(), hipMemcpyHostToDevice, s));
    dirty = false;
  }

  void step(std::vector<std::unique_ptr<GroupItem>> &finished) override {
    // retire finished runs
    for (size_t k = 0; k < runs.size();) {
      if (!runs[k]->R->iterating()) {
        runs[k]->R.reset();
        finished.push_back(std::move(runs[k]));
        runs.erase(runs.begin() + k);
        dirty = true;
      } else {
        k++;
      }
    }
    if (runs.empty()) return;
    if (dirty) rebuild();
    const int G = (int)runs.size();
    const RunSeg<T> *sg = d_segs.p;
    // the requests' preconditioners overlap on their own streams
    for (auto &g : runs) {
      g->S->precond(g->S->ri.p, g->S->zi.p);
      ELP_HIP(hipEventRecord(g->ready->e, g->own));
      ELP_HIP(hipStreamWaitEvent(s, g->ready->e, 0));
    }
    hipLaunchKernelGGL((grp_dot_kernel<T, B, 0>), dim3(r_blocks), dim3(256), 0, s, sg, G);
    hipLaunchKernelGGL((grp_final_kernel<T, B, 0>), dim3(G), dim3(kRedFinalThreads), 0, s, sg);
    hipLaunchKernelGGL((grp_beta_kernel<T, B>), dim3(G), dim3(64), 0, s, sg);
    hipLaunchKernelGGL((grp_dir_kernel<T, B>), dim3(e_blocks), dim3(256), 0, s, sg, G);
    ELP_HIP(hipGetLastError());
    // restart flags go up on the group's stream (the requests' streams are
    // idle until the next preconditioner)
    for (auto &g : runs) {
      g->R->s = s;
      g->R->restart_flags();
      g->R->s = g->own;
    }
    ELP_HIP(hipEventRecord(ev0.e, s));
    hipLaunchKernelGGL((grp_op_kernel<T, B>), dim3(op_blocks), dim3(256), 0, s, sg, G);
    ELP_HIP(hipEventRecord(ev1.e, s));
    hipLaunchKernelGGL((grp_dot_kernel<T, B, 1>), dim3(r_blocks), dim3(256), 0, s, sg, G);
    hipLaunchKernelGGL((grp_final_kernel<T, B, 1>), dim3(G), dim3(kRedFinalThreads), 0, s, sg);
    hipLaunchKernelGGL((grp_alpha_kernel<T, B>), dim3(G), dim3(64), 0, s, sg);
    hipLaunchKernelGGL((grp_update_kernel<T, B>), dim3(r_blocks), dim3(256), 0, s, sg, G);
    hipLaunchKernelGGL((grp_final_kernel<T, B, 2>), dim3(G), dim3(kRedFinalThreads), 0, s, sg);
    ELP_HIP(hipGetLastError());
    for (auto &g : runs) {
      Solver<T, B> &S = *g->S;
      S.u_of = S.fused_box() ? S.ri.p : nullptr;
      ELP_HIP(hipMemcpyAsync(S.h_red.p, S.red_out.p, sizeof(double2) * B, hipMemcpyDeviceToHost, s));
    }
    ELP_HIP(hipStreamSynchronize(s));
    float ms = 0;
    ELP_HIP(hipEventElapsedTime(&ms, ev0.e, ev1.e));
    for (auto &g : runs) {
      g->S->spmv_ms += ms;
      g->S->spmv_calls++;
      g->R->after_norms();
    }
  }
};

static int slot_index(int B) { return B == 4 ? 0 : B == 8 ? 1 : B == 16 ? 2 : B == 32 ? 3 : -1; }

template <typename T, int B>
static std::unique_ptr<GroupItem> make_group_run(Model &m, Model::SolverState &st, int64_t nf, const double *freqs,
                                                 std::complex<double> Z, const elp_solve_opts &o, double *X,
                                                 double *relres, int32_t *iters, hipStream_t s) {
  auto g = std::make_unique<GroupRun<T, B>>(m, freqs, nf, Z, o, X, relres, iters);
  g->S.reset(new Solver<T, B>(m, st, g->o, s));
  return g;
}

template <typename T>
static std::unique_ptr<GroupItem> make_group_run_b(int B, Model &m, Model::SolverState &st, int64_t nf,
                                                   const double *freqs, std::complex<double> Z,
                                                   const elp_solve_opts &o, double *X, double *relres,
                                                   int32_t *iters, hipStream_t s) {
  switch (B) {
    case 4: return make_group_run<T, 4>(m, st, nf, freqs, Z, o, X, relres, iters, s);
    case 8: return make_group_run<T, 8>(m, st, nf, freqs, Z, o, X, relres, iters, s);
    case 16: return make_group_run<T, 16>(m, st, nf, freqs, Z, o, X, relres, iters, s);
    default: return make_group_run<T, 32>(m, st, nf, freqs, Z, o, X, relres, iters, s);
  }
}

std::unique_ptr<GroupItem> group_prepare(Model &m, int64_t nf, const double *freqs, std::complex<double> Z,
                                         const elp_solve_opts &o, double *X, double *relres, int32_t *iters,
                                         hipStream_t setup, int *key) {
  ELP_REQUIRE(o.method == ELP_METHOD_COCG, ELP_EINVAL, "method must be ELP_METHOD_COCG");
  ELP_REQUIRE(o.rtol > 0 && o.max_iters > 0, ELP_EINVAL, "invalid tolerance / iteration limit");
  ELP_REQUIRE(operator_variant_valid(o.op_tile_rows), ELP_EINVAL, "op_tile_rows must be 0, -1, -2, -4, a tile size below 1000 or 2000/4000/6000 + a tile size");
  for (int64_t k = 0; k < nf; k++) ELP_REQUIRE(freqs[k] > 0, ELP_EINVAL, "frequencies must be positive");
  const int bi = slot_index(o.block);
  if (o.precision == ELP_PREC_FP64 || bi < 0) {  // one frequency at a time: runs alone
    *key = -1;
    return nullptr;
  }
  Model::SolverState &st = solver_state(m, o, Z, setup);
  const bool dbl = st.fp64_inner;  // coupled models iterate in FP64 (see sweep)
  // requests batch with others of their size class (row counts within a
  // factor of 4: floor(log4 n)), so that a small request's step is not the
  // length of a large one's
  int size_class = 0;
  for (int64_t v = m.n; v >= 4; v /= 4) size_class++;
  *key = (dbl ? 4 : 0) + bi + 8 * size_class;
  return dbl ? make_group_run_b<double>(o.block, m, st, nf, freqs, Z, o, X, relres, iters, setup)
             : make_group_run_b<float>(o.block, m, st, nf, freqs, Z, o, X, relres, iters, setup);
}

std::unique_ptr<GroupDriver> group_driver(int key, Ctx *ctx, hipStream_t s) {
  switch (key % 8) {
    case 0: return std::make_unique<GroupDriverT<float, 4>>(ctx, s);
    case 1: return std::make_unique<GroupDriverT<float, 8>>(ctx, s);
    case 2: return std::make_unique<GroupDriverT<float, 16>>(ctx, s);
    case 3: return std::make_unique<GroupDriverT<float, 32>>(ctx, s);
    case 4: return std::make_unique<GroupDriverT<double, 4>>(ctx, s);
    case 5: return std::make_unique<GroupDriverT<double, 8>>(ctx, s);
    case 6: return std::make_unique<GroupDriverT<double, 16>>(ctx, s);
    case 7: return std::make_unique<GroupDriverT<double, 32>>(ctx, s);
    default: throw Error(ELP_EINVAL, "unknown group");
  }
}

static const LevelData &L0b(const Model::SolverState &st) { return *st.levels[0]; }

template <typename T, int B>
static void run_sweep(Model &m, Model::SolverState &st, int64_t nf, const double *freqs, std::complex<double> Z,
                      const elp_solve_opts &o, double *X, double *relres, int32_t *iters, elp_sweep_stats *stats,
                      hipStream_t s) {
  if constexpr (std::is_same<T, float>::value && B == 32) {
    if (!st.op_chosen || st.op_tile_rows != o.op_tile_rows) {
      st.kept_solver.reset();  // its graph may hold the other variant's launches
      choose_operator(m, st, o, s);
    }
    if (!st.box_chosen && !st.fp64_inner) choose_box_post(m, st, s);
  }
  // small models keep their solver (work vectors, captured graph) for the
  // next sweep with the same options: their sweeps are bound by host work
  const int kind = (std::is_same<T, float>::value ? 0 : 64) + B;
  const bool keep = m.n <= kGraphAutoRows;
  std::unique_ptr<Solver<T, B>> own;
  Solver<T, B> *sp = nullptr;
  if (keep && st.kept_solver && st.kept_kind == kind && same_opts(st.kept_opts, o)) {
    own.reset(static_cast<Solver<T, B> *>(st.kept_solver.release()));  // back only if the sweep completes
    own->rearm(s);
  } else {
    st.kept_solver.reset();
    own.reset(new Solver<T, B>(m, st, o, s));
  }
  sp = own.get();
  Solver<T, B> &S = *sp;
  // launch-bound models replay each FP32 iteration from a captured graph
  S.use_graph = std::is_same<T, float>::value && (o.graph > 0 || (o.graph == 0 && m.n <= kGraphAutoRows));
  int refinements = 0;
  if (std::is_same<T, float>::value || B > 1) {
    solve_slots<T, B>(S, nf, freqs, Z, X, relres, iters, refinements);
  } else {
    for (int64_t f0 = 0; f0 < nf; f0 += B) {
      const int nb = (int)std::min<int64_t>(B, nf - f0);
      solve_block<T, B>(S, freqs + f0, nb, Z, X ? X + 2 * f0 * m.n : nullptr, relres ? relres + f0 : nullptr,
                        iters ? iters + f0 : nullptr, refinements);
    }
  }
  if (stats) {
    stats->spmv_calls = S.spmv_calls;
    stats->spmv_ms = S.spmv_ms;
    using V = typename VT<T>::t;
    const int64_t nx = (int64_t)m.x_col.size();
    stats->spmv_bytes = m.nnz * (int64_t)(4 + 2 * sizeof(T)) + (m.n + 1) * 4 * 2 + m.n +
                        nx * (int64_t)(4 + 1 + sizeof(T)) + 2 * m.n * B * (int64_t)sizeof(V);
    stats->refinements = refinements;
    stats->graph_replays = S.graph_replays;
    stats->box_tile_rows = (S.fused_box() && L0b(st).bt_on && o.op_tile_rows != -1) ? L0b(st).bt_rows : 0;
    stats->refactors = S.refactors;
    stats->refactor_ms = S.refactor_ms;
    const LevelData &L0 = *st.levels[0];
    stats->op_tile_rows = (std::is_same<T, float>::value && B == 32 && L0.pad_rp.p)
                              ? operator_variant(L0)
                              : 0;
  }
  if (keep) {
    st.kept_solver = std::unique_ptr<void, void (*)(void *)>(
        own.release(), [](void *p) { delete static_cast<Solver<T, B> *>(p); });
    st.kept_kind = kind;
    st.kept_opts = o;
  }
}

int64_t solver_bytes_estimate(const Model &m, int block, int precision) {
  // models with structure and fluid rows iterate in FP64; beside a Hex8
  // box they keep two factorizations of the structure block (the window in
  // use and the next): about 7 ns m complex FP64 values each with their work
  // space, m the half-bandwidth of the structure block, about 3 sqrt(ns)
  // for plates
  const bool cpl = coupled_possible(m);
  const int64_t vb = (precision == ELP_PREC_FP64 || two_fields(m)) ? 16 : 8;
  int64_t b = m.nnz * 8 + m.n * 24;                    // FP32 shadows, l1 sums
  if (cpl) {
    int64_t ns = 0;
    for (auto v : m.row_field) ns += v == 0;
    const double mb = std::min<double>((double)ns, 3.0 * std::sqrt((double)ns));
    b += (int64_t)(2 * 7.0 * (double)ns * mb * 16.0);
  }
  b += m.nnz * 18;                                     // permuted columns, row-padded operator copy
  b += m.n * 512;                                      // box products A P, padded P rows, lattice
  b += (int64_t)m.n * block * (5 * vb * 2 + 4 * 16);   // level-0 + Krylov + FP64 outer vectors
  if (precision != ELP_PREC_FP64) b += m.n * 16 * 16 * 2;  // narrow FP64 refinement blocks (x, A x)
  return b + b / 4;                                    // coarse levels (about 1/8 each) and slack
}

int64_t sweep(Model &m, int64_t nf, const double *freqs, std::complex<double> Z, const elp_solve_opts &o,
              double *X, double *relres_out, int32_t *iters, elp_sweep_stats *stats, hipStream_t s) {
  ELP_REQUIRE(o.method == ELP_METHOD_COCG, ELP_EINVAL, "method must be ELP_METHOD_COCG");
  std::vector<double> own_rel(relres_out ? 0 : nf);
  double *relres = relres_out ? relres_out : own_rel.data();
  ELP_REQUIRE(o.rtol > 0 && o.max_iters > 0, ELP_EINVAL, "invalid tolerance / iteration limit");
  ELP_REQUIRE(operator_variant_valid(o.op_tile_rows), ELP_EINVAL, "op_tile_rows must be 0, -1, -2, -4, a tile size below 1000 or 2000/4000/6000 + a tile size");
  for (int64_t k = 0; k < nf; k++) ELP_REQUIRE(freqs[k] > 0, ELP_EINVAL, "frequencies must be positive");
  auto t0 = std::chrono::steady_clock::now();
  Model::SolverState &st = solver_state(m, o, Z, s);
  auto t1 = std::chrono::steady_clock::now();
  const int B = o.precision == ELP_PREC_FP64 ? 1 : o.block;
  if (o.precision == ELP_PREC_FP64) {
    run_sweep<double, 1>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s);
  } else if (st.fp64_inner) {
    // Coupled models batch their frequencies like the mixed solver but keep
    // the inner Krylov in FP64: a thin plate's stiffness has cond ~1e8, so
    // the FP32 operator's own rounding (~6e-8 relative) leaves FP32 inner
    // solves that no FP64 refinement step can improve on, and the pressure
    // (held to rtol through the scaled residual) is driven by the plate
    switch (B) {
      case 1: run_sweep<double, 1>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      case 4: run_sweep<double, 4>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      case 8: run_sweep<double, 8>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      case 16: run_sweep<double, 16>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      case 32: run_sweep<double, 32>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      default: throw Error(ELP_EINVAL, "block must be 1, 4, 8, 16 or 32");
    }
  } else {
    switch (B) {
      case 1: run_sweep<float, 1>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      case 4: run_sweep<float, 4>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      case 8: run_sweep<float, 8>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      case 16: run_sweep<float, 16>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      case 32: run_sweep<float, 32>(m, st, nf, freqs, Z, o, X, relres, iters, stats, s); break;
      default: throw Error(ELP_EINVAL, "block must be 1, 4, 8, 16 or 32");
    }
  }
  auto t2 = std::chrono::steady_clock::now();
  int64_t unconverged = 0;
  for (int64_t k = 0; k < nf; k++) unconverged += !(relres[k] <= o.rtol);
  if (stats) {
    stats->setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    stats->solve_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    stats->unconverged = (int32_t)std::min<int64_t>(unconverged, INT32_MAX);
  }
  return unconverged;
}

}  // namespace elp
