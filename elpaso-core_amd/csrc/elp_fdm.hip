// Fast-diagonalization solve of the Hex8 box operator (see elp_fdm.h).
//
// 1-D uniform linear elements with natural (rigid-wall) ends are diagonalized
// by cosine modes: K v = lambda M v with v_m(i) = cos(pi m i / N),
// lambda_m = (6/h^2)(1 - cos t)/(2 + cos t), t = pi m / N.  The modes are
// normalized so V^T M V = I.
#include <hip/hip_runtime.h>

#include <cmath>

#include "elp_fdm.h"

namespace elp {


struct FdmBox {
  int dims[3];  // px, py, pz (nodes per direction)
  double h[3];
  double rho, c;
  int robin_lo, robin_hi;  // Robin on z = z0 / z = z1
  DArr<float> fx32, ix32, fy32, iy32;   // forward (V^T) and inverse (V) transforms
  DArr<double> fx64, ix64, fy64, iy64;
  DArr<double> lamx, lamy;
};

static void fdm_free(FdmBox *f) { delete f; }

int64_t fdm_nodes(const FdmBox &f) { return (int64_t)f.dims[0] * f.dims[1] * f.dims[2]; }
const int *fdm_dims(const FdmBox &f) { return f.dims; }
double fdm_rho(const FdmBox &f) { return f.rho; }
double fdm_c(const FdmBox &f) { return f.c; }

// M-orthonormal cosine modes of a uniform 1-D mesh of n elements of size h
static void modes_1d(int n, double h, std::vector<double> &V, std::vector<double> &lam) {
  const int p = n + 1;
  V.assign((size_t)p * p, 0.0);
  lam.assign(p, 0.0);
  for (int m = 0; m < p; m++) {
    const double t = M_PI * m / n;
    lam[m] = 6.0 / (h * h) * (1.0 - std::cos(t)) / (2.0 + std::cos(t));
    std::vector<double> v(p);
    for (int i = 0; i < p; i++) v[i] = std::cos(t * i);
    // v^T M v with the consistent mass (h/6)[2 1; 1 2] per element
    double nrm = 0.0;
    for (int e = 0; e < n; e++) nrm += h / 6.0 * (2 * v[e] * v[e] + 2 * v[e] * v[e + 1] + 2 * v[e + 1] * v[e + 1]);
    const double s = 1.0 / std::sqrt(nrm);
    for (int i = 0; i < p; i++) V[(size_t)i * p + m] = v[i] * s;  // V[i][m]
  }
}

std::unique_ptr<FdmBox, void (*)(FdmBox *)> fdm_build(const Model &m, bool fluid_part) {
  return fdm_build(m, m.box, fluid_part ? 1 : 0);
}

std::unique_ptr<FdmBox, void (*)(FdmBox *)> fdm_build(const Model &m, const BoxInfo &bx, int mode) {
  std::unique_ptr<FdmBox, void (*)(FdmBox *)> none(nullptr, fdm_free);
  if (!bx.set) return none;
  // the box's fluid: the model's one fluid material
  const Material *fluid = nullptr;
  int n_fluid = 0;
  for (auto &kv : m.mats)
    if (kv.second.fluid) {
      fluid = &kv.second;
      n_fluid++;
    }
  if (n_fluid != 1 || (mode == 0 && m.mats.size() != 1)) return none;
  if (mode == 0) {
    // fluid rows, and rows of prescribed pressures (left out of the
    // correction by the interpolation)
    for (auto v : m.row_field)
      if (v == 0) return none;
  } else if (mode == 1) {
    // the box lattice covers exactly the fluid rows (a constrained fluid row
    // is not a fluid row: the count then differs)
    int64_t nf = 0;
    for (auto v : m.row_field) nf += v == 1;
    const int64_t lattice = (int64_t)(bx.nel[0] + 1) * (bx.nel[1] + 1) * (bx.nel[2] + 1);
    if (bx.order != 1 || nf != lattice) return none;
    for (auto &kb : m.kinds)
      if (is_fluid(kb.kind) && kb.kind != ELP_FLUID8) return none;
  }
  if (bx.robin_sides & ~((1 << 2) | (1 << 4))) return none;  // Robin only on z sides
  if (bx.nel[0] + 1 > kFdmMaxAxis || bx.nel[1] + 1 > kFdmMaxAxis) return none;
  // the uniform lattice of the box (the model's rows are mapped onto it by
  // the solver's interpolation)
  auto f = std::unique_ptr<FdmBox, void (*)(FdmBox *)>(new FdmBox(), fdm_free);
  for (int d = 0; d < 3; d++) {
    f->dims[d] = bx.nel[d] + 1;
    f->h[d] = bx.length[d] / bx.nel[d];
  }
  f->rho = fluid->rho;
  f->c = fluid->c;
  f->robin_lo = (bx.robin_sides >> 2) & 1;
  f->robin_hi = (bx.robin_sides >> 4) & 1;
  Ctx *c = m.ctx;
  auto upload = [&](int d, DArr<float> &fw32, DArr<float> &iv32, DArr<double> &fw64, DArr<double> &iv64,
                    DArr<double> *lam) {
    std::vector<double> V, L;
    modes_1d(bx.nel[d], f->h[d], V, L);
    const int p = f->dims[d];
    std::vector<double> Vt((size_t)p * p);
    for (int i = 0; i < p; i++)
      for (int k = 0; k < p; k++) Vt[(size_t)k * p + i] = V[(size_t)i * p + k];
    iv64.upload(c, V);
    fw64.upload(c, Vt);
    std::vector<float> a(V.begin(), V.end()), b(Vt.begin(), Vt.end());
    iv32.upload(c, a);
    fw32.upload(c, b);
    if (lam) lam->upload(c, L);
  };
  upload(0, f->fx32, f->ix32, f->fx64, f->ix64, &f->lamx);
  upload(1, f->fy32, f->iy32, f->fy64, f->iy64, &f->lamy);
  return f;
}

// out[o][p][c] = sum_l Mat[p][l] in[o][l][c].  A workgroup owns CW "columns"
// (o, c): a tile c0 .. c0+CW of one line o when inner >= CW, or
// floor(CW / inner) whole lines when inner < CW.  The columns' values for
// all l are staged in LDS once; Mat is staged by chunks of KC columns,
// transposed.  Thread (pg, cg) accumulates the 4 columns 4cg .. 4cg+3 for the
// NP modes p = pg, pg + PG, ... in registers, so each LDS read of the tile
// feeds 4 NP FMAs.  Every output sums over l in increasing order.
constexpr int kAxisKC = 16;
template <typename T, int CW, int NP>
__global__ __launch_bounds__(256) void axis_transform_kernel(int64_t outer, int L, int64_t inner,
                                                             const T *__restrict__ Mat, const T *__restrict__ in,
                                                             T *__restrict__ out) {
  constexpr int CG = CW / 4, PG = 256 / CG;
  extern __shared__ unsigned char smem[];
  T *tile = reinterpret_cast<T *>(smem);  // [L][CW]
  T *mt = tile + (size_t)L * CW;          // [KC][L]: mt[k][p] = Mat[p][l0 + k]
  // column j of this workgroup -> offset of (o, l = 0, c), or -1 past the end
  auto colbase = [&](int j) -> int64_t {
    int64_t o, c;
    if (inner >= CW) {
      o = blockIdx.x;
      c = (int64_t)blockIdx.y * CW + j;
      if (c >= inner) return -1;
    } else {
      const int lines = CW / (int)inner;
      if (j >= lines * inner) return -1;
      o = (int64_t)blockIdx.x * lines + j / inner;
      c = j % inner;
      if (o >= outer) return -1;
    }
    return o * L * inner + c;
  };
  for (int idx = threadIdx.x; idx < L * CW; idx += 256) {
    const int l = idx / CW, cc = idx % CW;
    const int64_t bse = colbase(cc);
    tile[idx] = bse >= 0 ? in[bse + (int64_t)l * inner] : T(0);
  }
  const int cg = threadIdx.x % CG, pg = threadIdx.x / CG;
  T acc[NP][4];
#pragma unroll
  for (int i = 0; i < NP; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0;
  for (int l0 = 0; l0 < L; l0 += kAxisKC) {
    const int kc = min(kAxisKC, L - l0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < kc * L; idx += 256) {
      const int p = idx / kc, k = idx % kc;
      mt[k * L + p] = Mat[(int64_t)p * L + l0 + k];
    }
    __syncthreads();
    for (int k = 0; k < kc; k++) {
      const T *tr = tile + (l0 + k) * CW + 4 * cg;
      const T t0 = tr[0], t1 = tr[1], t2 = tr[2], t3 = tr[3];
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const int p = pg + i * PG;
        const T mv = p < L ? mt[k * L + p] : T(0);
        acc[i][0] += mv * t0;
        acc[i][1] += mv * t1;
        acc[i][2] += mv * t2;
        acc[i][3] += mv * t3;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int64_t bse = colbase(4 * cg + j);
    if (bse < 0) continue;
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const int p = pg + i * PG;
      if (p < L) out[bse + (int64_t)p * inner] = acc[i][j];
    }
  }
}

// Transforms on the matrix cores, out[o][p][c] = sum_l Mat[p][l] in[o][l][c]
// (v_mfma_f32_16x16x4f32 for FP32, v_mfma_f64_16x16x4f64 for FP64): one
// persistent 512-thread workgroup per CU stages Mat once in LDS (rows padded
// with zeros to RB 16, columns to a multiple of 4, row stride KP + 1), and
// each wavefront walks strips of 16 columns (o, c0 .. c0+15) on its own:
// lane j + 16 k loads the B fragment element in[o][4 ks + k][c0 + j] straight
// from memory, CH reduction steps ahead of the matrix cores (the next
// strip's first steps while the current one finishes), and accumulates the
// RB row blocks of the strip.  Fragment layouts (checked on the device): A
// element (i, k) from lane i + 16 k, B element (k, j) from lane j + 16 k;
// the result (4 (l / 16) + r, l % 16) in element r of lane l for FP32,
// (l / 16 + 4 r, l % 16) for FP64.  Each product adds its 4 terms in order
// and the padded terms are exact zeros: the same bits as
// axis_transform_kernel and as the earlier FP32 kernel that staged 64-column
// tiles in LDS beside Mat.  Measured per transform: FP64 on the 101 x 101 x
// 121 lattice of a 1.2 m cavity with 32 frequencies (630 MB per vector)
// 0.44 ms against 1.21 ms on the vector units; FP32 on the 101 x 81 x 61
// lattice of the Hex20 cavity (128 MB) 0.14 / 0.11 ms (x / y) against
// 0.20 / 0.16 ms with LDS-staged tiles.
template <typename T> struct MfmaStrip;
template <> struct MfmaStrip<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static constexpr int kMaxAxis = 160;  // RB <= 10: 160 x 161 floats of LDS
  static constexpr int kChunk = 8;
  __device__ static acc_t mfma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int k, int r) { return 4 * k + r; }
};
template <> struct MfmaStrip<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static constexpr int kMaxAxis = 128;  // RB <= 8: 128 x 129 doubles of LDS
  static constexpr int kChunk = 16;
  __device__ static acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int k, int r) { return k + 4 * r; }
};
template <typename T, int RB>
__global__ __launch_bounds__(512) void axis_strip_kernel(int64_t outer, int L, int64_t inner, int64_t nstrips,
                                                         const T *__restrict__ Mat, const T *__restrict__ in,
                                                         T *__restrict__ out) {
  using MS = MfmaStrip<T>;
  using acc_t = typename MS::acc_t;
  constexpr int CH = MS::kChunk;
  extern __shared__ unsigned char smem_strip[];
  T *sA = reinterpret_cast<T *>(smem_strip);
  const int KP = (L + 3) / 4 * 4, S = KP + 1;
  for (int idx = threadIdx.x; idx < RB * 16 * KP; idx += 512) {
    const int p = idx / KP, l = idx % KP;
    sA[p * S + l] = (p < L && l < L) ? Mat[(int64_t)p * L + l] : T(0);
  }
  __syncthreads();
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int i = lane % 16, k = lane / 16;
  const int64_t spl = inner / 16;
  const int nks = KP / 4, nch = (nks + CH - 1) / CH;
  const int64_t stride = (int64_t)gridDim.x * 8;
  // B fragments of chunk c (reduction steps CH c .. CH c + CH - 1) of strip
  // st; 0 past the last strip or the last row
  auto load = [&](int64_t st, int c, T *bb) {
    const int64_t o = st / spl, c0 = (st % spl) * 16;
    const T *src = in + o * L * inner + c0 + i;
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const int l = (c * CH + u) * 4 + k;
      bb[u] = (st < nstrips && l < L) ? src[(int64_t)l * inner] : T(0);
    }
  };
  int64_t strip = (int64_t)blockIdx.x * 8 + wave;
  int ch = 0;
  T bc[CH], bn[CH];
  load(strip, 0, bc);
  acc_t acc[RB];
#pragma unroll
  for (int pb = 0; pb < RB; pb++) acc[pb] = acc_t{0, 0, 0, 0};
  while (strip < nstrips) {
    int64_t ns = strip;
    int nc = ch + 1;
    if (nc == nch) {
      nc = 0;
      ns = strip + stride;
    }
    load(ns, nc, bn);
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const int ks = ch * CH + u;
      if (ks < nks) {
#pragma unroll
        for (int pb = 0; pb < RB; pb++) acc[pb] = MS::mfma(sA[(pb * 16 + i) * S + ks * 4 + k], bc[u], acc[pb]);
      }
    }
    if (ch == nch - 1) {
      const int64_t o = strip / spl, c0 = (strip % spl) * 16;
      T *dst = out + o * L * inner + c0 + i;
#pragma unroll
      for (int pb = 0; pb < RB; pb++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int p = pb * 16 + MS::row(k, r);
          if (p < L) dst[(int64_t)p * inner] = acc[pb][r];
        }
        acc[pb] = acc_t{0, 0, 0, 0};
      }
    }
#pragma unroll
    for (int u = 0; u < CH; u++) bc[u] = bn[u];
    strip = ns;
    ch = nc;
  }
}

template <typename V>
__device__ inline V cdiv(V a, V b) {
  const auto d = b.x * b.x + b.y * b.y;
  V r;
  r.x = (a.x * b.x + a.y * b.y) / d;
  r.y = (a.y * b.x - a.x * b.y) / d;
  return r;
}

// one complex tridiagonal z-system per (x-mode p, y-mode q, frequency f):
//   T = (alpha/w^2 - sM) Mz + Kz/w^2 + tG Ez,  alpha = lamx[p] + lamy[q];
// x = rho T^{-1} r (Thomas elimination; c' kept in cs)
template <typename T, typename V>
__global__ void zsolve_kernel(int px, int py, int pz, int B, double hz, int rlo, int rhi,
                              const FdmCoef *__restrict__ coef, const double *__restrict__ lamx,
                              const double *__restrict__ lamy, const V *__restrict__ r, V *__restrict__ x,
                              V *__restrict__ cs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t plane = (int64_t)px * py * B;
  if (t >= plane) return;
  const int f = (int)(t % B);
  const int64_t pq = t / B;
  const int p = (int)(pq % px), q = (int)(pq / px);
  const FdmCoef cf = coef[f];
  const double a = (lamx[p] + lamy[q]) * cf.w2inv;
  // coefficient of Mz: (a - sM) complex; of Kz: w2inv real
  const double cmr = a - cf.sM.x, cmi = -cf.sM.y;
  const double kd = 1.0 / hz, md = hz / 6.0;
  // off-diagonal (same for all k): (cm * md) + w2inv * (-kd)
  V off;
  off.x = (T)(cmr * md - cf.w2inv * kd);
  off.y = (T)(cmi * md);
  V cprev = {0, 0}, dprev = {0, 0};
  for (int k = 0; k < pz; k++) {
    const bool end = (k == 0 || k == pz - 1);
    const double mk = md * (end ? 2.0 : 4.0), kk = kd * (end ? 1.0 : 2.0);
    V diag;
    diag.x = (T)(cmr * mk + cf.w2inv * kk);
    diag.y = (T)(cmi * mk);
    if ((k == 0 && rlo) || (k == pz - 1 && rhi)) {
      diag.x += (T)cf.tG.x;
      diag.y += (T)cf.tG.y;
    }
    V m = diag;
    if (k > 0) {
      // m = diag - off * c'
      m.x -= off.x * cprev.x - off.y * cprev.y;
      m.y -= off.x * cprev.y + off.y * cprev.x;
    }
    V rv = r[k * plane + t];
    if (k > 0) {
      rv.x -= off.x * dprev.x - off.y * dprev.y;
      rv.y -= off.x * dprev.y + off.y * dprev.x;
    }
    const V cn = cdiv(off, m), dn = cdiv(rv, m);
    cs[k * plane + t] = cn;
    x[k * plane + t] = dn;
    cprev = cn;
    dprev = dn;
  }
  V nxt = x[(int64_t)(pz - 1) * plane + t];
  const T rho = (T)cf.rho;
  {
    V o = nxt;
    o.x *= rho;
    o.y *= rho;
    x[(int64_t)(pz - 1) * plane + t] = o;
  }
  for (int k = pz - 2; k >= 0; k--) {
    const V ck = cs[k * plane + t];
    V d = x[k * plane + t];
    d.x -= ck.x * nxt.x - ck.y * nxt.y;
    d.y -= ck.x * nxt.y + ck.y * nxt.x;
    nxt = d;
    V o = d;
    o.x *= rho;
    o.y *= rho;
    x[k * plane + t] = o;
  }
}

template <typename T, int CW, int NP>
static void launch_axis(int64_t outer, int L, int64_t inner, const T *Mat, const T *in, T *out, hipStream_t s) {
  const size_t lds = ((size_t)L * CW + (size_t)kAxisKC * L) * sizeof(T);
  const auto kern = axis_transform_kernel<T, CW, NP>;
  ELP_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const dim3 grid = inner >= CW ? dim3((unsigned)outer, (unsigned)((inner + CW - 1) / CW))
                                 : dim3((unsigned)((outer + CW / inner - 1) / (CW / inner)), 1);  // CW/inner lines
  hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, outer, L, inner, Mat, in, out);
  ELP_HIP(hipGetLastError());
}

template <typename T, int RB>
static void launch_axis_strip(int64_t outer, int L, int64_t inner, const T *Mat, const T *in, T *out, hipStream_t s) {
  const int KP = (L + 3) / 4 * 4;
  const size_t lds = (size_t)RB * 16 * (KP + 1) * sizeof(T);
  const int64_t nstrips = outer * (inner / 16);
  const auto kern = axis_strip_kernel<T, RB>;
  ELP_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const unsigned grid = (unsigned)std::min<int64_t>((nstrips + 7) / 8, 256);  // one workgroup per CU
  hipLaunchKernelGGL(kern, dim3(grid), dim3(512), lds, s, outer, L, inner, nstrips, Mat, in, out);
  ELP_HIP(hipGetLastError());
}


// transform along an axis of L <= kFdmMaxAxis points; inner counts reals
template <typename T, typename V>
static void transform(int64_t outer, int L, int64_t inner, const T *Mat, const V *in, V *out, hipStream_t s) {
  if (L <= MfmaStrip<T>::kMaxAxis && inner % 16 == 0) {  // whole 16-column strips (B >= 8 frequencies)
    const T *i = reinterpret_cast<const T *>(in);
    T *o = reinterpret_cast<T *>(out);
    switch ((L + 15) / 16) {
      case 1: launch_axis_strip<T, 1>(outer, L, inner, Mat, i, o, s); break;
      case 2: launch_axis_strip<T, 2>(outer, L, inner, Mat, i, o, s); break;
      case 3: launch_axis_strip<T, 3>(outer, L, inner, Mat, i, o, s); break;
      case 4: launch_axis_strip<T, 4>(outer, L, inner, Mat, i, o, s); break;
      case 5: launch_axis_strip<T, 5>(outer, L, inner, Mat, i, o, s); break;
      case 6: launch_axis_strip<T, 6>(outer, L, inner, Mat, i, o, s); break;
      case 7: launch_axis_strip<T, 7>(outer, L, inner, Mat, i, o, s); break;
      case 8: launch_axis_strip<T, 8>(outer, L, inner, Mat, i, o, s); break;
      case 9: launch_axis_strip<T, std::is_same<T, float>::value ? 9 : 8>(outer, L, inner, Mat, i, o, s); break;
      default: launch_axis_strip<T, std::is_same<T, float>::value ? 10 : 8>(outer, L, inner, Mat, i, o, s); break;
    }
    return;
  }
  constexpr int CW = sizeof(T) == 4 ? 64 : 32;
  ELP_REQUIRE(L <= kFdmMaxAxis && inner > 0, ELP_EINVAL, "box transform shape");
  constexpr int PG = 256 / (CW / 4);
  const T *i = reinterpret_cast<const T *>(in);
  T *o = reinterpret_cast<T *>(out);
  const int np = (L + PG - 1) / PG;
  if (np <= 2) launch_axis<T, CW, 2>(outer, L, inner, Mat, i, o, s);
  else if (np <= 4) launch_axis<T, CW, 4>(outer, L, inner, Mat, i, o, s);
  else if (np <= 8) launch_axis<T, CW, 8>(outer, L, inner, Mat, i, o, s);
  else launch_axis<T, CW, 16>(outer, L, inner, Mat, i, o, s);
}

template <typename T, typename V>
static void solve_impl(const FdmBox &f, int B, const FdmCoef *coef, const V *b, V *x, V *s1, V *s2, const T *fx,
                       const T *ix, const T *fy, const T *iy, hipStream_t s) {
  const int px = f.dims[0], py = f.dims[1], pz = f.dims[2];
  // x-modes: mix the fastest index (blocks of px x B complex)
  transform<T, V>((int64_t)py * pz, px, 2LL * B, fx, b, s1, s);
  // y-modes: per z-plane, mix j over rows of px x B complex
  transform<T, V>((int64_t)pz, py, 2LL * B * px, fy, s1, s2, s);
  const int64_t plane = (int64_t)px * py * B;
  hipLaunchKernelGGL((zsolve_kernel<T, V>), dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, px, py, pz, B,
                     f.h[2], f.robin_lo, f.robin_hi, coef, f.lamx.p, f.lamy.p, s2, s1, x);
  ELP_HIP(hipGetLastError());
  transform<T, V>((int64_t)pz, py, 2LL * B * px, iy, s1, s2, s);
  transform<T, V>((int64_t)py * pz, px, 2LL * B, ix, s2, x, s);
}

void fdm_solve(const FdmBox &f, int B, const FdmCoef *coef, const float2 *b, float2 *x, float2 *s1, float2 *s2,
               hipStream_t s) {
  solve_impl<float, float2>(f, B, coef, b, x, s1, s2, f.fx32.p, f.ix32.p, f.fy32.p, f.iy32.p, s);
}
void fdm_solve(const FdmBox &f, int B, const FdmCoef *coef, const double2 *b, double2 *x, double2 *s1, double2 *s2,
               hipStream_t s) {
  solve_impl<double, double2>(f, B, coef, b, x, s1, s2, f.fx64.p, f.ix64.p, f.fy64.p, f.iy64.p, s);
}

}  // namespace elp
