// Host symbolic phase of a model: elPaSo's DOF numbering, the CSR sparsity
// pattern, the node -> element adjacency that drives the deterministic gather,
// the compact face/FSI entries, loads and constraints; then upload.
//
// Reference behaviour restated here (paths relative to elpasoCore/source):
//  - DOF numbering: cAnalysis::activateDofsAtNodes (analysis/analysis.cpp:475-520):
//    nodes in ascending id, DOFs of a node by enum value, consecutive rows.
//  - Pattern: getLocalSparsityPattern + preallocateMatrix (analysis.cpp:783-867):
//    element DOF cliques, interface blocks C and C^T, and the diagonal.
//  - Element map order (ascending element id) is the order in which elPaSo
//    adds element matrices (analysisfrequencybasic.cpp:281-304); the
//    adjacency lists keep it so the device sums in the same order.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>

#include "elp_fem.h"
#include "elp_internal.h"

namespace elp {

namespace {

constexpr uint32_t kStructMask = 0x7u;        // disp_x1..disp_x3
constexpr uint32_t kFluidMask = 1u << 16;     // fluid

uint32_t kind_mask(int kind) { return is_fluid(kind) ? kFluidMask : kStructMask; }

// The box description (origin, lengths, elements per axis, order, Robin
// sides) read from the mesh when the caller gave none, so that meshes read
// from files get the box preconditioners too.  Found only when the fluid
// elements are of one kind and tile a uniform axis-aligned lattice exactly:
// each element occupies its own cell with its nodes at the cell's lattice
// points (corners, and for Hex20 the edge midpoints), and every Robin face
// lies on a side of the box.  Anything else leaves the box unset.
void detect_box(Model &m) {
  if (m.box.set) return;
  const KindBlock *fk = nullptr;
  for (auto &kb : m.kinds)
    if (is_fluid(kb.kind)) {
      if (fk) return;  // Fluid8 and Fluid20 together
      fk = &kb;
    }
  if (!fk || fk->n_elem == 0) return;
  const int order = fk->kind == ELP_FLUID20 ? 2 : 1;
  const int npe = fk->npe;
  const int(*nat)[3] = npe == 8 ? kHex8Nat : kHex20Nat;
  auto X = [&](int32_t r, int d) { return m.xyz[(int64_t)m.rank_to_input[r] * 3 + d]; };
  // the lattice coordinates of each axis: the distinct values over the
  // fluid nodes, evenly spaced
  double lo[3], hi[3], h[3];
  int64_t nel[3];
  for (int d = 0; d < 3; d++) {
    std::vector<double> v;
    v.reserve(fk->conn.size());
    for (int32_t r : fk->conn) v.push_back(X(r, d));
    std::sort(v.begin(), v.end());
    lo[d] = v.front();
    hi[d] = v.back();
    const double ext = hi[d] - lo[d];
    if (!(ext > 0)) return;
    const double tol = 1e-9 * ext;
    std::vector<double> u;
    for (double x : v)
      if (u.empty() || x - u.back() > tol) u.push_back(x);
    const int64_t p = (int64_t)u.size();
    if ((p - 1) % order) return;
    nel[d] = (p - 1) / order;
    const double step = ext / (double)(p - 1);
    for (int64_t k = 0; k < p; k++)
      if (std::fabs(u[k] - (lo[d] + step * (double)k)) > tol) return;
    h[d] = ext / (double)nel[d];
  }
  if (nel[0] * nel[1] * nel[2] != fk->n_elem) return;
  // each element on its own cell, its nodes at the cell's lattice points
  std::vector<uint8_t> taken((size_t)fk->n_elem, 0);
  for (int64_t e = 0; e < fk->n_elem; e++) {
    const int32_t *c = &fk->conn[e * npe];
    int64_t cell[3];
    for (int d = 0; d < 3; d++) {
      double mn = X(c[0], d);
      for (int a = 1; a < npe; a++) mn = std::min(mn, X(c[a], d));
      cell[d] = (int64_t)std::llround((mn - lo[d]) / h[d]);
      if (cell[d] < 0 || cell[d] >= nel[d]) return;
      for (int a = 0; a < npe; a++) {
        const double want = lo[d] + h[d] * ((double)cell[d] + 0.5 * (double)(nat[a][d] + 1));
        if (std::fabs(X(c[a], d) - want) > 1e-9 * (hi[d] - lo[d])) return;
      }
    }
    uint8_t &t = taken[(size_t)((cell[2] * nel[1] + cell[1]) * nel[0] + cell[0])];
    if (t) return;
    t = 1;
  }
  // Robin faces on the box sides (bit k: 0-based Hex8 face k's normal,
  // -y, +y, -z, +x, +z, -x)
  static const int side_bit[3][2] = {{5, 3}, {0, 1}, {2, 4}};
  int sides = 0;
  for (int64_t f = 0; f < m.n_faces; f++) {
    const int32_t *fn = &m.face_nodes[f * m.nfn];
    int bit = -1;
    for (int d = 0; d < 3 && bit < 0; d++)
      for (int s = 0; s < 2 && bit < 0; s++) {
        const double at = s ? hi[d] : lo[d];
        bool all = true;
        for (int k = 0; k < m.nfn; k++) all &= std::fabs(X(fn[k], d) - at) <= 1e-9 * (hi[d] - lo[d]);
        if (all) bit = side_bit[d][s];
      }
    if (bit < 0) return;  // a Robin face inside the box
    sides |= 1 << bit;
  }
  for (int d = 0; d < 3; d++) {
    m.box.origin[d] = lo[d];
    m.box.length[d] = hi[d] - lo[d];
    m.box.nel[d] = (int)nel[d];
  }
  m.box.order = order;
  m.box.robin_sides = sides;
  m.box.set = true;
  m.box.detected = true;
}

}  // namespace

void finalize_model(Model &m) {
  ELP_REQUIRE(!m.finalized, ELP_ESTATE, "model already finalized");
  ELP_REQUIRE(!m.node_ids.empty(), ELP_EINVAL, "model has no nodes");
  ELP_REQUIRE(!m.groups.empty(), ELP_EINVAL, "model has no elements");
  const int64_t N = (int64_t)m.node_ids.size();
  ELP_REQUIRE(N < (1ll << 31) - 1, ELP_EINVAL, "too many nodes for 32-bit indices");
  m.n_nodes = N;

  // ---- node ranks (ascending id)
  std::vector<int32_t> ord(N);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return m.node_ids[a] < m.node_ids[b]; });
  m.sorted_ids.resize(N);
  m.rank_to_input = ord;
  for (int64_t r = 0; r < N; r++) m.sorted_ids[r] = m.node_ids[ord[r]];
  for (int64_t r = 1; r < N; r++)
    ELP_REQUIRE(m.sorted_ids[r] != m.sorted_ids[r - 1], ELP_EINVAL,
                "duplicate node id " + std::to_string(m.sorted_ids[r]));
  auto rank_of = [&](int64_t id) -> int32_t {
    auto it = std::lower_bound(m.sorted_ids.begin(), m.sorted_ids.end(), id);
    if (it == m.sorted_ids.end() || *it != id) throw Error(ELP_EINVAL, "unknown node id " + std::to_string(id));
    return (int32_t)(it - m.sorted_ids.begin());
  };

  // ---- elements in ascending id across all groups
  struct ERef { int64_t id; int32_t g; int64_t k; };
  std::vector<ERef> all;
  for (int32_t g = 0; g < (int32_t)m.groups.size(); g++) {
    auto &G = m.groups[g];
    ELP_REQUIRE(m.mats.count(G.mat), ELP_EINVAL, "element group references unknown material " + std::to_string(G.mat));
    const Material &mat = m.mats[G.mat];
    ELP_REQUIRE(mat.fluid == is_fluid(G.kind), ELP_EINVAL, "material type does not match element kind");
    for (int64_t k = 0; k < (int64_t)G.ids.size(); k++) all.push_back({G.ids[k], g, k});
  }
  std::sort(all.begin(), all.end(), [](const ERef &a, const ERef &b) { return a.id < b.id; });
  for (size_t i = 1; i < all.size(); i++)
    ELP_REQUIRE(all[i].id != all[i - 1].id, ELP_EINVAL, "duplicate element id " + std::to_string(all[i].id));

  // ---- kind blocks: element order within a kind = ascending id
  int kind_slot[5] = {-1, -1, -1, -1, -1};
  for (auto &e : all) {
    int kind = m.groups[e.g].kind;
    if (kind_slot[kind] < 0) {
      kind_slot[kind] = (int)m.kinds.size();
      m.kinds.emplace_back();
      m.kinds.back().kind = kind;
      m.kinds.back().npe = npe_of(kind);
      m.kinds.back().nd = nd_of(kind);
    }
  }
  std::vector<uint32_t> mask(N, 0);
  std::vector<int8_t> node_kind(N, -1);
  // global element order -> (kind slot, kind-local index)
  for (auto &e : all) {
    auto &G = m.groups[e.g];
    KindBlock &kb = m.kinds[kind_slot[G.kind]];
    const Material &mat = m.mats[G.mat];
    const int npe = kb.npe;
    for (int a = 0; a < npe; a++) {
      int32_t r = rank_of(G.conn[e.k * npe + a]);
      kb.conn.push_back(r);
      mask[r] |= kind_mask(G.kind);
      if (node_kind[r] < 0) node_kind[r] = (int8_t)kind_slot[G.kind];
      ELP_REQUIRE(node_kind[r] == kind_slot[G.kind], ELP_EINVAL,
                  "node " + std::to_string(m.sorted_ids[r]) + " is shared by elements of different kinds");
    }
    if (is_fluid(G.kind)) {
      kb.params.insert(kb.params.end(), {mat.rho, 1.0 / (mat.c * mat.c), mat.rho, mat.c});
    } else {
      // isotropic D: lambda = f nu, mu = f (1-2nu)/2, f = E/(1+nu)/(1-2nu)
      // (materialstructureisotrop.cpp:130-147)
      double f = mat.E / (1.0 + mat.nu) / (1.0 - 2.0 * mat.nu);
      kb.params.insert(kb.params.end(), {f * mat.nu, f * (1.0 - 2.0 * mat.nu) / 2.0, mat.rho, mat.E});
    }
    kb.n_elem++;
  }
  // visco: one loss factor per model (applied as a per-frequency scalar)
  bool any_struct = false;
  for (auto &kv : m.mats) {
    if (kv.second.fluid) continue;
    if (!any_struct) {
      m.eta = kv.second.eta;
      m.visco_type = kv.second.visco_type;
      any_struct = true;
    } else {
      ELP_REQUIRE(kv.second.eta == m.eta && (m.eta == 0 || kv.second.visco_type == m.visco_type), ELP_EINVAL,
                  "all structural materials of a model must share the loss factor and visco type");
    }
  }
  m.has_visco = any_struct && m.eta != 0.0;

  // ---- DOF numbering (activateDofsAtNodes)
  m.node_row.assign(N, -1);
  m.node_nd.assign(N, 0);
  m.node_dofmask = mask;
  int64_t row = 0;
  for (int64_t r = 0; r < N; r++) {
    if (!mask[r]) continue;
    m.node_row[r] = (int32_t)row;
    int nd = __builtin_popcount(mask[r]);
    m.node_nd[r] = nd;
    for (int d = 0; d < nd; d++) m.row_field.push_back(mask[r] == kFluidMask ? 1 : 0);
    row += nd;
  }
  ELP_REQUIRE(row < (1ll << 31) - 1, ELP_EINVAL, "too many DOFs for 32-bit indices");
  m.n = row;

  // ---- adjacency per kind (node -> elements, ascending element id)
  for (auto &kb : m.kinds) {
    std::vector<int32_t> cnt(N, 0);
    for (int32_t r : kb.conn) cnt[r]++;
    std::vector<int32_t> slot(N, -1);
    for (int64_t r = 0; r < N; r++)
      if (cnt[r]) {
        slot[r] = (int32_t)kb.nodes.size();
        kb.nodes.push_back((int32_t)r);
      }
    kb.adj_ptr.assign(kb.nodes.size() + 1, 0);
    for (size_t i = 0; i < kb.nodes.size(); i++) kb.adj_ptr[i + 1] = kb.adj_ptr[i] + cnt[kb.nodes[i]];
    kb.adj_elem.resize(kb.adj_ptr.back());
    kb.adj_local.resize(kb.adj_ptr.back());
    std::vector<int32_t> fill(kb.nodes.size(), 0);
    for (int64_t e = 0; e < kb.n_elem; e++)
      for (int a = 0; a < kb.npe; a++) {
        int32_t s = slot[kb.conn[e * kb.npe + a]];
        int32_t p = kb.adj_ptr[s] + fill[s]++;
        kb.adj_elem[p] = (int32_t)e;
        kb.adj_local[p] = (uint8_t)a;
      }
  }

  // ---- interface elements (sorted by id)
  m.n_if = (int64_t)m.if_ids.size();
  std::vector<int32_t> if_sranks;
  if (m.n_if) {
    std::vector<int64_t> io(m.n_if);
    std::iota(io.begin(), io.end(), 0);
    std::sort(io.begin(), io.end(), [&](int64_t a, int64_t b) { return m.if_ids[a] < m.if_ids[b]; });
    for (int64_t e : io) {
      for (int k = 0; k < 4; k++) {
        int32_t fr = rank_of(m.if_fn[e * 4 + k]), sr = rank_of(m.if_sn[e * 4 + k]);
        ELP_REQUIRE(mask[fr] == kFluidMask, ELP_EINVAL, "interface fluid node without a fluid DOF");
        ELP_REQUIRE(mask[sr] == kStructMask, ELP_EINVAL, "interface structure node without displacement DOFs");
        m.if_fnodes.push_back(fr);
        if_sranks.push_back(sr);
      }
      m.if_factor.push_back(-1.0 * m.if_ori[e]);  // factorC = factorCT = -ori
    }
  }

  // ---- node graph: element neighbours + interface partners
  std::vector<std::vector<int32_t>> extra_nbr;  // only for interface nodes
  std::vector<int32_t> extra_ptr;
  {
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (int64_t e = 0; e < m.n_if; e++)
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          pairs.push_back({if_sranks[e * 4 + a], m.if_fnodes[e * 4 + b]});
          pairs.push_back({m.if_fnodes[e * 4 + b], if_sranks[e * 4 + a]});
        }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    extra_ptr.assign(N + 1, 0);
    for (auto &p : pairs) extra_ptr[p.first + 1]++;
    for (int64_t r = 0; r < N; r++) extra_ptr[r + 1] += extra_ptr[r];
    extra_nbr.resize(1);
    extra_nbr[0].resize(pairs.size());
    for (size_t i = 0; i < pairs.size(); i++) extra_nbr[0][i] = pairs[i].second;  // sorted by first
  }
  // per node: the kind block and its slot
  std::vector<int32_t> node_slot(N, -1);
  for (auto &kb : m.kinds)
    for (size_t i = 0; i < kb.nodes.size(); i++) node_slot[kb.nodes[i]] = (int32_t)i;
  std::vector<int32_t> deg(N, 0);
  std::vector<std::vector<int32_t>> lists((size_t)N);
  parallel_for(N, [&](int64_t a, int64_t b) {
    std::vector<int32_t> tmp;
    for (int64_t r = a; r < b; r++) {
      if (node_kind[r] < 0) continue;
      const KindBlock &kb = m.kinds[node_kind[r]];
      int32_t s = node_slot[r];
      tmp.clear();
      for (int32_t p = kb.adj_ptr[s]; p < kb.adj_ptr[s + 1]; p++) {
        const int32_t *c = &kb.conn[(int64_t)kb.adj_elem[p] * kb.npe];
        tmp.insert(tmp.end(), c, c + kb.npe);
      }
      for (int32_t p = extra_ptr[r]; p < extra_ptr[r + 1]; p++) tmp.push_back(extra_nbr[0][p]);
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      lists[r] = tmp;
      deg[r] = (int32_t)tmp.size();
    }
  });
  m.nbr_ptr.assign(N + 1, 0);
  for (int64_t r = 0; r < N; r++) m.nbr_ptr[r + 1] = m.nbr_ptr[r] + deg[r];
  m.nbr.resize(m.nbr_ptr.back());
  m.nbr_off.resize(m.nbr_ptr.back());
  std::vector<int32_t> rowlen(N, 0);
  for (int64_t r = 0; r < N; r++) {
    int32_t off = 0;
    for (int32_t i = 0; i < deg[r]; i++) {
      int32_t j = lists[r][i];
      m.nbr[m.nbr_ptr[r] + i] = j;
      m.nbr_off[m.nbr_ptr[r] + i] = off;
      off += m.node_nd[j];
    }
    rowlen[r] = off;
    std::vector<int32_t>().swap(lists[r]);
  }
  // the assembly kernels address a node's neighbours by one-byte slots
  // (rows themselves may be of any length)
  for (int64_t r = 0; r < N; r++)
    ELP_REQUIRE(deg[r] <= 256, ELP_EINVAL,
                "node " + std::to_string(m.sorted_ids[r]) + " is coupled to more than 256 nodes");

  // ---- CSR
  m.rowptr.assign(m.n + 1, 0);
  int64_t nnz = 0;
  for (int64_t r = 0; r < N; r++)
    for (int d = 0; d < m.node_nd[r]; d++) {
      nnz += rowlen[r];
      m.rowptr[m.node_row[r] + d + 1] = 0;
    }
  ELP_REQUIRE(nnz < (1ll << 31) - 1, ELP_EINVAL, "more than 2^31 nonzeros");
  m.nnz = nnz;
  {
    int64_t acc = 0;
    for (int64_t r = 0; r < N; r++)
      for (int d = 0; d < m.node_nd[r]; d++) {
        m.rowptr[m.node_row[r] + d] = (int32_t)acc;
        acc += rowlen[r];
      }
    m.rowptr[m.n] = (int32_t)acc;
  }
  m.colind.resize(nnz);
  parallel_for(N, [&](int64_t a, int64_t b) {
    for (int64_t r = a; r < b; r++) {
      for (int d = 0; d < m.node_nd[r]; d++) {
        int64_t p = m.rowptr[m.node_row[r] + d];
        for (int32_t i = m.nbr_ptr[r]; i < m.nbr_ptr[r + 1]; i++) {
          int32_t j = m.nbr[i];
          for (int dd = 0; dd < m.node_nd[j]; dd++) m.colind[p++] = m.node_row[j] + dd;
        }
      }
    }
  });
  // ---- records of the fused Hex8 kernel (see KindBlock::h8_desc)
  for (auto &kb : m.kinds) {
    if (kb.npe != 8) continue;
    const int64_t ns = (int64_t)kb.nodes.size();
    const int32_t slots_per_pass = kb.nd == 1 ? 64 : 32;
    // simple: at most 8 elements, one pass of slots, every neighbour a node of
    // the same field (no interface partner), so slot t is columns nd t.. of the rows
    std::vector<char> simple(ns);
    parallel_for(ns, [&](int64_t a, int64_t b) {
      for (int64_t s = a; s < b; s++) {
        const int32_t i = kb.nodes[s];
        bool ok = kb.adj_ptr[s + 1] - kb.adj_ptr[s] <= 8 && m.nbr_ptr[i + 1] - m.nbr_ptr[i] <= slots_per_pass;
        for (int32_t p = m.nbr_ptr[i]; ok && p < m.nbr_ptr[i + 1]; p++) ok = m.node_nd[m.nbr[p]] == kb.nd;
        simple[s] = ok;
      }
    });
    std::vector<int64_t> at(ns);  // descriptor position: simple nodes first, in node order
    kb.h8_simple = 0;
    for (int64_t s = 0; s < ns; s++)
      if (simple[s]) at[s] = kb.h8_simple++;
    for (int64_t s = 0, o = kb.h8_simple; s < ns; s++)
      if (!simple[s]) at[s] = o++;
    kb.h8_desc.assign(ns * 8, 0);
    kb.adj_slot.assign(kb.adj_elem.size() * 8, 0);
    parallel_for(ns, [&](int64_t a, int64_t b) {
      for (int64_t s = a; s < b; s++) {
        const int32_t i = kb.nodes[s], r0 = m.node_row[i];
        int32_t *d = &kb.h8_desc[at[s] * 8];
        d[0] = kb.adj_ptr[s];
        d[1] = kb.adj_ptr[s + 1] - kb.adj_ptr[s];
        d[2] = m.nbr_ptr[i];
        d[3] = m.nbr_ptr[i + 1] - m.nbr_ptr[i];
        d[4] = m.rowptr[r0];
        d[5] = m.rowptr[r0 + 1] - m.rowptr[r0];
        const int32_t *nb = m.nbr.data() + m.nbr_ptr[i];
        for (int32_t p = d[0]; p < d[0] + d[1]; p++)
          for (int q = 0; q < 8; q++) {
            const int32_t j = kb.conn[(int64_t)kb.adj_elem[p] * 8 + q];
            kb.adj_slot[(int64_t)p * 8 + q] = (uint8_t)(std::lower_bound(nb, nb + d[3], j) - nb);
          }
      }
    });
  }
  // ---- 20-node kinds: neighbour slots per adjacency entry and the record of
  // each (element, local node) (see KindBlock::row_at)
  for (auto &kb : m.kinds) {
    if (kb.npe != 20) continue;
    const int64_t ns = (int64_t)kb.nodes.size();
    kb.adj_slot.assign(kb.adj_elem.size() * 20, 0);
    kb.row_at.assign(kb.n_elem * 20, -1);
    parallel_for(ns, [&](int64_t a, int64_t b) {
      for (int64_t s = a; s < b; s++) {
        const int32_t i = kb.nodes[s];
        const int32_t *nb = m.nbr.data() + m.nbr_ptr[i];
        const int32_t dg = m.nbr_ptr[i + 1] - m.nbr_ptr[i];
        for (int32_t p = kb.adj_ptr[s]; p < kb.adj_ptr[s + 1]; p++) {
          const int64_t e = kb.adj_elem[p];
          kb.row_at[e * 20 + kb.adj_local[p]] = p;  // each (element, local node) once
          for (int q = 0; q < 20; q++) {
            const int32_t j = kb.conn[e * 20 + q];
            kb.adj_slot[(int64_t)p * 20 + q] = (uint8_t)(std::lower_bound(nb, nb + dg, j) - nb);
          }
        }
      }
    });
    for (int32_t v : kb.row_at) ELP_REQUIRE(v >= 0, ELP_EINVAL, "element node without adjacency entry");
  }
  auto full_pos = [&](int32_t r, int32_t c) -> int32_t {
    auto b = m.colind.begin() + m.rowptr[r], e = m.colind.begin() + m.rowptr[r + 1];
    auto it = std::lower_bound(b, e, c);
    if (it == e || *it != c) throw Error(ELP_EINVAL, "entry outside the sparsity pattern");
    return (int32_t)(it - m.colind.begin());
  };

  // ---- Robin faces (element loads: elements in id order, faces ascending)
  const HexTables &T = host_tables();
  struct XE { int32_t row, col; uint8_t type; int32_t idx; int64_t order; };
  std::vector<XE> xe;
  int64_t order = 0;
  if (!m.imp_elem.empty()) {
    std::map<int64_t, std::pair<int, int64_t>> eid;  // id -> (kind slot, local index)
    {
      std::vector<int64_t> cnt(m.kinds.size(), 0);
      for (auto &e : all) {
        int ks = kind_slot[m.groups[e.g].kind];
        eid[e.id] = {ks, cnt[ks]++};
      }
    }
    std::vector<std::pair<int64_t, int>> fl;
    for (size_t i = 0; i < m.imp_elem.size(); i++) fl.push_back({m.imp_elem[i], (int)m.imp_face[i]});
    std::sort(fl.begin(), fl.end());
    fl.erase(std::unique(fl.begin(), fl.end()), fl.end());
    for (auto &f : fl) {
      auto it = eid.find(f.first);
      ELP_REQUIRE(it != eid.end(), ELP_EINVAL, "impedance face on unknown element " + std::to_string(f.first));
      const KindBlock &kb = m.kinds[it->second.first];
      ELP_REQUIRE(is_fluid(kb.kind), ELP_EINVAL, "impedance face on a structural element");
      ELP_REQUIRE(f.second >= 1 && f.second <= 6, ELP_EINVAL, "face number must be 1..6");
      int nfn = kb.npe == 8 ? 4 : 8;
      ELP_REQUIRE(m.nfn == 0 || m.nfn == nfn, ELP_EINVAL, "impedance faces of mixed order");
      m.nfn = nfn;
      const int32_t *c = &kb.conn[it->second.second * kb.npe];
      for (int k = 0; k < nfn; k++)
        m.face_nodes.push_back(c[kb.npe == 8 ? T.face8[f.second - 1][k] : T.face20[f.second - 1][k]]);
    }
    m.n_faces = (int64_t)fl.size();
    const int nfn = m.nfn;
    for (int64_t f = 0; f < m.n_faces; f++) {
      const int32_t *fn = &m.face_nodes[f * nfn];
      for (int z = 0; z < nfn; z++)
        for (int s = 0; s < nfn; s++) {
          int32_t rz = m.node_row[fn[z]], rs = m.node_row[fn[s]];
          xe.push_back({rz, m.imp_lumped ? rz : rs, 0, (int32_t)(f * nfn * nfn + z * nfn + s), order++});
        }
    }
  }
  // ---- FSI entries: C (12 x 4) into (structure rows, fluid cols), C^T back
  const int64_t if_base = m.n_faces * m.nfn * m.nfn;
  for (int64_t e = 0; e < m.n_if; e++) {
    for (int z = 0; z < 12; z++)
      for (int s = 0; s < 4; s++) {
        int32_t sr = m.node_row[if_sranks[e * 4 + z / 3]] + z % 3;
        int32_t fc = m.node_row[m.if_fnodes[e * 4 + s]];
        xe.push_back({sr, fc, 1, (int32_t)(if_base + e * 48 + z * 4 + s), order++});
      }
    for (int z = 0; z < 4; z++)
      for (int s = 0; s < 12; s++) {
        int32_t fr = m.node_row[m.if_fnodes[e * 4 + z]];
        int32_t sc = m.node_row[if_sranks[e * 4 + s / 3]] + s % 3;
        xe.push_back({fr, sc, 1, (int32_t)(if_base + e * 48 + s * 4 + z), order++});
      }
  }
  std::stable_sort(xe.begin(), xe.end(), [](const XE &a, const XE &b) {
    return a.row != b.row ? a.row < b.row : (a.col != b.col ? a.col < b.col : a.type < b.type);
  });
  m.x_rowptr.assign(m.n + 1, 0);
  m.xc_ptr.push_back(0);
  for (size_t i = 0; i < xe.size(); i++) {
    bool fresh = i == 0 || xe[i].row != xe[i - 1].row || xe[i].col != xe[i - 1].col || xe[i].type != xe[i - 1].type;
    if (fresh) {
      if (i) m.xc_ptr.push_back((int32_t)m.xc_idx.size());
      m.x_col.push_back(xe[i].col);
      m.x_type.push_back(xe[i].type);
      m.x_fullpos.push_back(full_pos(xe[i].row, xe[i].col));
      m.x_rowptr[xe[i].row + 1]++;
    }
    m.xc_idx.push_back(xe[i].idx);
  }
  if (!xe.empty()) m.xc_ptr.push_back((int32_t)m.xc_idx.size());
  for (int64_t r = 0; r < m.n; r++) m.x_rowptr[r + 1] += m.x_rowptr[r];

  // ---- loads (applyNodalForces, analysis.cpp:171-279)
  m.F.assign(m.n, {0.0, 0.0});
  for (auto &s : m.sources) {
    int32_t r = rank_of(s.first);
    ELP_REQUIRE(mask[r] == kFluidMask, ELP_EINVAL, "point source on a node without fluid DOF");
    m.F[m.node_row[r]] += 4.0 * M_PI * s.second * 1.0 * 1.0;
  }
  for (auto &f : m.forces) {
    int32_t r = rank_of(f.first);
    ELP_REQUIRE(mask[r] == kStructMask, ELP_EINVAL, "point force on a node without displacement DOFs");
    for (int d = 0; d < 3; d++) m.F[m.node_row[r] + d] += f.second[d];
  }
  // ---- Dirichlet (getGlobalBCs / insertBoundaryConditions, analysis.cpp:366-454)
  {
    std::map<int32_t, std::complex<double>> dm;
    for (auto &d : m.dirichlet) {
      int32_t r = rank_of(d.node);
      ELP_REQUIRE(d.dof >= 0 && d.dof < 21, ELP_EINVAL, "invalid dof");
      if (!(mask[r] & (1u << d.dof))) continue;  // only active DOFs are constrained
      int32_t row = m.node_row[r] + __builtin_popcount(mask[r] & ((1u << d.dof) - 1));
      dm[row] = d.v;
    }
    for (auto &kv : dm) {
      m.dir_rows.push_back(kv.first);
      m.dir_vals.push_back(kv.second);
      m.row_field[kv.first] = 2;
    }
  }

  // ---- upload
  Ctx *c = m.ctx;
  std::vector<double> xyz_r(N * 3);
  for (int64_t r = 0; r < N; r++)
    for (int d = 0; d < 3; d++) xyz_r[r * 3 + d] = m.xyz[(int64_t)m.rank_to_input[r] * 3 + d];
  m.d_xyz.upload(c, xyz_r);
  m.d_rowptr.upload(c, m.rowptr);
  m.d_colind.upload(c, m.colind);
  m.d_nbr_ptr.upload(c, m.nbr_ptr);
  m.d_nbr.upload(c, m.nbr);
  m.d_nbr_off.upload(c, m.nbr_off);
  m.d_node_row.upload(c, m.node_row);
  m.d_row_field.upload(c, m.row_field);
  for (auto &kb : m.kinds) {
    kb.d_conn.upload(c, kb.conn);
    kb.d_nodes.upload(c, kb.nodes);
    kb.d_adj_ptr.upload(c, kb.adj_ptr);
    kb.d_adj_elem.upload(c, kb.adj_elem);
    kb.d_adj_local.upload(c, kb.adj_local);
    kb.d_h8_desc.upload(c, kb.h8_desc);
    kb.d_adj_slot.upload(c, kb.adj_slot);
    kb.d_row_at.upload(c, kb.row_at);
    kb.d_params.upload(c, kb.params);
  }
  m.d_x_rowptr.upload(c, m.x_rowptr);
  m.d_x_col.upload(c, m.x_col);
  m.d_x_type.upload(c, m.x_type);
  m.d_x_fullpos.upload(c, m.x_fullpos);
  m.d_xc_ptr.upload(c, m.xc_ptr);
  m.d_xc_idx.upload(c, m.xc_idx);
  m.d_face_nodes.upload(c, m.face_nodes);
  m.d_if_fnodes.upload(c, m.if_fnodes);
  m.d_if_factor.upload(c, m.if_factor);
  post_prepare(m);
  detect_box(m);
  m.finalized = true;
}

int64_t model_device_bytes_estimate(const Model &m) {
  int64_t b = 0;
  b += m.nnz * (4 + 8 * 2 + (m.has_rayleigh ? 8 : 0));
  b += (m.n + 1) * 4 * 2;
  for (auto &kb : m.kinds) {
    int64_t nd = (int64_t)kb.nd * kb.npe;
    b += kb.n_elem * (nd * nd + (int64_t)kb.npe * kb.npe) * 8;
    if (kb.npe == 20) b += kb.n_elem * 20 * (20 + 4);  // neighbour slots, record positions
  }
  return b;
}

}  // namespace elp
