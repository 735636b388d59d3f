// Internal C++ types of libelpaso_amd: context, device buffers, the model
// (host description + finalized device layout) and error handling.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/elpaso_amd.h"

namespace elp {

// f(a, b) over [0, n) split into contiguous chunks on up to 16 host threads
template <typename F>
void parallel_for(int64_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    int64_t a = t * chunk, b = std::min<int64_t>(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=]() { f(a, b); });
  }
  for (auto &x : th) x.join();
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define ELP_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::elp::Error(ELP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define ELP_REQUIRE(cond, code, msg)                   \
  do {                                                 \
    if (!(cond)) throw ::elp::Error((code), (msg));    \
  } while (0)

void set_last_error(const std::string &msg);

struct Ctx {
  int device = 0;
  std::atomic<int64_t> bytes{0};  // device bytes held by models of this context
  hipStream_t stream = nullptr;   // default work stream of the context
  ~Ctx() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// RAII HIP event (destroyed also when an error unwinds the caller).
struct HipEvent {
  hipEvent_t e = nullptr;
  HipEvent() { ELP_HIP(hipEventCreate(&e)); }
  HipEvent(const HipEvent &) = delete;
  HipEvent &operator=(const HipEvent &) = delete;
  ~HipEvent() {
    if (e) (void)hipEventDestroy(e);
  }
};

// RAII pinned host array (hipHostMalloc).
template <typename T>
struct PinnedArr {
  T *p = nullptr;
  PinnedArr() = default;
  PinnedArr(const PinnedArr &) = delete;
  PinnedArr &operator=(const PinnedArr &) = delete;
  void alloc(size_t n) {
    ELP_REQUIRE(!p, ELP_ESTATE, "pinned array allocated twice");
    ELP_HIP(hipHostMalloc((void **)&p, sizeof(T) * n, hipHostMallocDefault));
  }
  T &operator[](size_t i) { return p[i]; }
  const T &operator[](size_t i) const { return p[i]; }
  ~PinnedArr() {
    if (p) (void)hipHostFree(p);
  }
};

// RAII device allocation, accounted against the owning context.
template <typename T>
struct DArr {
  T *p = nullptr;
  size_t n = 0;
  Ctx *ctx = nullptr;
  DArr() = default;
  DArr(const DArr &) = delete;
  DArr &operator=(const DArr &) = delete;
  DArr(DArr &&o) noexcept { *this = std::move(o); }
  DArr &operator=(DArr &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n; ctx = o.ctx;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DArr() { release(); }
  void alloc(Ctx *c, size_t count) {
    release();
    ctx = c;
    n = count;
    if (count) {
      ELP_HIP(hipMalloc(&p, count * sizeof(T)));
      if (ctx) ctx->bytes += (int64_t)(count * sizeof(T));
    }
  }
  void upload(Ctx *c, const std::vector<T> &h) {
    alloc(c, h.size());
    if (n) ELP_HIP(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
  void download(std::vector<T> &h) const {
    h.resize(n);
    if (n) ELP_HIP(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
  }
  void zero(hipStream_t s = nullptr) {
    if (n) ELP_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  size_t bytes() const { return n * sizeof(T); }
  void release() {
    if (p) {
      (void)hipFree(p);
      if (ctx) ctx->bytes -= (int64_t)(n * sizeof(T));
    }
    p = nullptr;
    n = 0;
  }
};

struct Material {
  bool fluid = true;
  double rho = 0, c = 0, E = 0, nu = 0, eta = 0;
  int visco_type = 1;
};

struct ElemGroupIn {
  int kind;
  int64_t mat;
  std::vector<int64_t> ids, conn;
};

inline int npe_of(int kind) { return (kind == ELP_FLUID8 || kind == ELP_BRICK8) ? 8 : 20; }
inline int nd_of(int kind) { return (kind == ELP_FLUID8 || kind == ELP_FLUID20) ? 1 : 3; }
inline bool is_fluid(int kind) { return kind == ELP_FLUID8 || kind == ELP_FLUID20; }

// One kind of volume element after finalize: elements of the kind in
// ascending id, their node ranks, per-element material parameters, and the
// node-side adjacency used by the deterministic gather.
struct KindBlock {
  int kind = 0, npe = 0, nd = 0;
  int64_t n_elem = 0;
  std::vector<int32_t> conn;      // n_elem x npe node ranks
  std::vector<double> params;     // n_elem x 4 (fluid: rho, 1/c^2, rho, c ; brick: lambda, mu, rho, E)
  std::vector<int32_t> nodes;     // node ranks whose adjacent elements are of this kind
  std::vector<int32_t> adj_ptr;   // per entry of `nodes`: adjacent elements (kind-local, ascending id)
  std::vector<int32_t> adj_elem;
  std::vector<uint8_t> adj_local; // local node index in the element
  // fused Hex8 kernel: per entry of `nodes` {first adjacency entry, adjacent
  // elements, first neighbour, neighbours, first CSR entry of the node's
  // first row, row length, 0, 0}; per adjacency entry the neighbour slot of
  // each of the element's 8 nodes
  // Nodes in at most 8 elements with at most 32 (3 DOF) or 64 (1 DOF)
  // neighbour slots, all of them nodes of the node's own field, come first,
  // h8_simple of them, then the others.
  std::vector<int32_t> h8_desc;
  std::vector<uint8_t> adj_slot;
  int64_t h8_simple = 0;
  // 20-node kinds: adj_slot holds the neighbour slot of each of the element's
  // 20 nodes per adjacency entry, and row_at (per element and local node) the
  // adjacency entry whose record receives that node's rows of the element
  // matrices (records in node order, so the gather reads them contiguously)
  std::vector<int32_t> row_at;
  DArr<int32_t> d_row_at;
  // fluid kinds: runs of whole entries of `nodes` for the node-results
  // gradient kernel (block b: entries grad_blk[b] .. grad_blk[b + 1]); empty
  // if a node is in too many elements for one block
  std::vector<int32_t> grad_blk;
  DArr<int32_t> d_grad_blk;
  DArr<int32_t> d_conn, d_nodes, d_adj_ptr, d_adj_elem, d_h8_desc;
  DArr<uint8_t> d_adj_local, d_adj_slot;
  DArr<double> d_params;
  DArr<double> d_Ke, d_Me;        // element matrices (nd npe)^2 and npe^2 per element
};

struct BoxInfo {
  bool set = false;
  bool detected = false;  // read from the mesh at finalize (no hint given)
  double origin[3], length[3];
  int nel[3], order = 1, robin_sides = 0;
};

struct Model {
  // keeps the context alive while any device array of the model exists
  // (declared first so it is destroyed last)
  std::shared_ptr<Ctx> ctx_ref;
  Ctx *ctx = nullptr;
  // ---- input
  std::vector<int64_t> node_ids;
  std::vector<double> xyz;
  std::map<int64_t, Material> mats;
  std::vector<ElemGroupIn> groups;
  std::vector<int64_t> imp_elem, imp_face;
  std::complex<double> Z{0, 0};
  bool imp_lumped = true;
  std::vector<int64_t> if_ids, if_fn, if_sn;
  std::vector<double> if_ori;
  std::vector<std::pair<int64_t, double>> sources;
  std::vector<std::pair<int64_t, std::array<double, 3>>> forces;
  struct Dir { int64_t node; int dof; std::complex<double> v; };
  std::vector<Dir> dirichlet;
  double ray_a = 0, ray_b = 0;
  bool has_rayleigh = false;
  BoxInfo box;
  // ---- finalized
  bool finalized = false, assembled = false;
  int64_t n_nodes = 0, n = 0, nnz = 0;
  std::vector<int64_t> sorted_ids;       // node ids ascending (rank -> id)
  std::vector<int32_t> rank_to_input;    // rank -> position in input arrays
  std::vector<int32_t> node_row, node_nd; // first row and #dofs per node rank
  std::vector<uint32_t> node_dofmask;
  std::vector<uint8_t> row_field;        // 0 structure, 1 fluid, 2 Dirichlet
  std::vector<int32_t> rowptr, colind;
  std::vector<int32_t> nbr_ptr, nbr, nbr_off;  // node graph, column offset of each neighbour
  std::vector<KindBlock> kinds;
  // extra entries (C_Gamma and FSI S) in a compact CSR over all rows
  std::vector<int32_t> x_rowptr, x_col, x_fullpos;
  std::vector<uint8_t> x_type;           // 0 = C_Gamma, 1 = S
  std::vector<int32_t> xc_ptr, xc_idx;   // contributions into the face/FSI buffer
  std::vector<int32_t> face_nodes;       // n_faces x nfn node ranks (face node order)
  int nfn = 0;
  int64_t n_faces = 0, n_if = 0;
  std::vector<int32_t> if_fnodes;        // n_if x 4 fluid node ranks
  std::vector<double> if_factor;         // -ori per interface
  // visco
  double eta = 0;
  int visco_type = 1;
  bool has_visco = false;
  // rhs and Dirichlet
  std::vector<std::complex<double>> F;   // without Dirichlet values
  std::vector<int32_t> dir_rows;
  std::vector<std::complex<double>> dir_vals;
  // device
  DArr<double> d_xyz;
  DArr<int32_t> d_rowptr, d_colind, d_nbr_ptr, d_nbr, d_nbr_off, d_node_row;
  DArr<double> d_K, d_M, d_C;
  DArr<int32_t> d_x_rowptr, d_x_col, d_x_fullpos, d_xc_ptr, d_xc_idx, d_face_nodes, d_if_fnodes;
  DArr<uint8_t> d_x_type, d_row_field;
  DArr<double> d_x_val, d_fbuf, d_if_factor;
  // node results: node -> (kind slot, entry of the kind's node list) and the
  // shape-derivative tables (fluid kinds at their nodes, bricks at the centroid)
  DArr<int8_t> d_post_slot;
  DArr<int32_t> d_post_pos;
  DArr<double> d_T8, d_T20, d_Tc;
  // solver-side state (built lazily)
  struct SolverState;
  std::shared_ptr<SolverState> solver;
  std::mutex mu;
};

// host-side builders (elp_model.cpp)
void finalize_model(Model &m);
int64_t model_device_bytes_estimate(const Model &m);
// assembly driver (elp_assembly.hip)
void assemble_model(Model &m, hipStream_t s);
void upload_tables(Ctx *ctx);
float time_volume_assembly(Model &m, int repeats, hipStream_t s);
void element_matrices_batch(Ctx *ctx, int kind, int64_t n, const double *xyz, double p0,
                            double p1, double p2, double *Ke, double *Me);
void dynamic_matrix(Model &m, double omega, std::complex<double> Z, double *a, double *f);
// element stresses (bricks) and node results (sigma, grad p) of solutions X
void node_results(Model &m, int64_t nf, const double *freqs, const double *X, double *cells, double *nodes,
                  double *kernel_ms);
// node results of device-resident solutions into device memory, with the
// library's static per-device scratch instead of allocations
void node_results_noalloc(Model &m, int64_t nf, const double *freqs, const double *X_dev, double *nodes_dev);
void post_prepare(Model &m);          // per-model tables of the node results (at finalize)
void post_scratch_init(int device);   // address of the static scratch on `device` (current)
int64_t post_scratch_bytes();
int64_t brick_element_count(const Model &m);

}  // namespace elp

// opaque C-ABI handles
struct elp_ctx {
  std::shared_ptr<elp::Ctx> c;
};
struct elp_model {
  std::unique_ptr<elp::Model> m;
};

