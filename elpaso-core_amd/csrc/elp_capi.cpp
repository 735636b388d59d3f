// extern "C" entry points of libelpaso_amd (declared in include/elpaso_amd.h).
// Every entry point catches all C++ exceptions and turns them into a status
// code plus a thread-local message (elp_last_error).
#include <chrono>
#include <cmath>

#include "elp_internal.h"
#include "elp_solver.h"


namespace elp {
static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace elp

#define ELP_TRY try {
#define ELP_CATCH                                     \
  }                                                   \
  catch (const elp::Error &e) {                       \
    elp::set_last_error(e.what());                    \
    return e.code;                                    \
  }                                                   \
  catch (const std::bad_alloc &) {                    \
    elp::set_last_error("host out of memory");        \
    return ELP_EMEM;                                  \
  }                                                   \
  catch (const std::exception &e) {                   \
    elp::set_last_error(e.what());                    \
    return ELP_EINVAL;                                \
  }                                                   \
  return ELP_OK;

#define ELP_ARG(cond, msg) ELP_REQUIRE(cond, ELP_EINVAL, msg)

static elp::Model &M(elp_model *m) {
  ELP_ARG(m && m->m, "null model");
  return *m->m;
}

extern "C" {

int elp_ctx_create(elp_ctx **out, int device) {
  ELP_TRY
  ELP_ARG(out, "null output pointer");
  int n = 0;
  ELP_HIP(hipGetDeviceCount(&n));
  ELP_ARG(device >= 0 && device < n, "invalid device index");
  ELP_HIP(hipSetDevice(device));
  auto c = std::make_shared<elp::Ctx>();
  c->device = device;
  ELP_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  elp::upload_tables(c.get());
  elp::post_scratch_init(device);
  auto *h = new elp_ctx();
  h->c = c;
  *out = h;
  ELP_CATCH
}

void elp_ctx_destroy(elp_ctx *ctx) {
  // models created from the context keep it alive until they are destroyed
  delete ctx;
}

const char *elp_last_error(const elp_ctx *) { return elp::g_last_error.c_str(); }

int elp_default_solve_opts(elp_solve_opts *o) {
  ELP_TRY
  ELP_ARG(o, "null opts");
  elp::default_opts(*o);
  ELP_CATCH
}

int elp_model_create(elp_ctx *ctx, elp_model **out) {
  ELP_TRY
  ELP_ARG(ctx && out, "null argument");
  auto *m = new elp_model();
  m->m.reset(new elp::Model());
  m->m->ctx_ref = ctx->c;
  m->m->ctx = ctx->c.get();
  *out = m;
  ELP_CATCH
}

void elp_model_destroy(elp_model *m) {
  if (!m) return;
  if (m->m && m->m->ctx) (void)hipSetDevice(m->m->ctx->device);
  delete m;
}

int elp_model_set_nodes(elp_model *mm, int64_t n, const int64_t *ids, const double *xyz) {
  ELP_TRY
  auto &m = M(mm);
  ELP_REQUIRE(!m.finalized, ELP_ESTATE, "model already finalized");
  ELP_ARG(n > 0 && ids && xyz, "invalid nodes");
  m.node_ids.assign(ids, ids + n);
  m.xyz.assign(xyz, xyz + 3 * n);
  ELP_CATCH
}

int elp_model_add_fluid_material(elp_model *mm, int64_t id, double rho, double c) {
  ELP_TRY
  auto &m = M(mm);
  ELP_ARG(rho > 0 && c > 0, "fluid material needs rho > 0 and c > 0");
  elp::Material mat;
  mat.fluid = true;
  mat.rho = rho;
  mat.c = c;
  m.mats[id] = mat;
  ELP_CATCH
}

int elp_model_add_structure_material(elp_model *mm, int64_t id, double E, double nu, double rho, double eta,
                                     int visco_type) {
  ELP_TRY
  auto &m = M(mm);
  ELP_ARG(E > 0 && rho > 0 && nu > -1.0 && nu < 0.5, "invalid structural material");
  ELP_ARG(eta >= 0 && (visco_type == 1 || visco_type == 2 || (visco_type == 0 && eta == 0)),
          "invalid loss factor / visco type");
  elp::Material mat;
  mat.fluid = false;
  mat.E = E;
  mat.nu = nu;
  mat.rho = rho;
  mat.eta = eta;
  mat.visco_type = visco_type == 0 ? 1 : visco_type;  // 0: elastic (eta = 0)
  m.mats[id] = mat;
  ELP_CATCH
}

int elp_model_add_elements(elp_model *mm, int kind, int64_t material_id, int64_t n, const int64_t *elem_ids,
                           const int64_t *conn) {
  ELP_TRY
  auto &m = M(mm);
  ELP_REQUIRE(!m.finalized, ELP_ESTATE, "model already finalized");
  ELP_ARG(kind >= ELP_FLUID8 && kind <= ELP_BRICK20, "unknown element kind");
  ELP_ARG(n >= 0 && (n == 0 || (elem_ids && conn)), "invalid elements");
  elp::ElemGroupIn g;
  g.kind = kind;
  g.mat = material_id;
  g.ids.assign(elem_ids, elem_ids + n);
  g.conn.assign(conn, conn + n * elp::npe_of(kind));
  m.groups.push_back(std::move(g));
  ELP_CATCH
}

int elp_model_add_impedance_faces(elp_model *mm, int64_t n, const int64_t *elem_ids, const int32_t *faces,
                                  double z_re, double z_im, int lumped) {
  ELP_TRY
  auto &m = M(mm);
  ELP_REQUIRE(!m.finalized, ELP_ESTATE, "model already finalized");
  ELP_ARG(n >= 0 && (n == 0 || (elem_ids && faces)), "invalid faces");
  for (int64_t i = 0; i < n; i++) {
    m.imp_elem.push_back(elem_ids[i]);
    m.imp_face.push_back(faces[i]);
  }
  m.Z = {z_re, z_im};
  m.imp_lumped = lumped != 0;
  ELP_CATCH
}

int elp_model_add_interfaces(elp_model *mm, int64_t n, const int64_t *ids, const double *ori,
                             const int64_t *fluid_nodes, const int64_t *struct_nodes) {
  ELP_TRY
  auto &m = M(mm);
  ELP_REQUIRE(!m.finalized, ELP_ESTATE, "model already finalized");
  ELP_ARG(n >= 0 && (n == 0 || (ids && ori && fluid_nodes && struct_nodes)), "invalid interfaces");
  m.if_ids.insert(m.if_ids.end(), ids, ids + n);
  m.if_ori.insert(m.if_ori.end(), ori, ori + n);
  m.if_fn.insert(m.if_fn.end(), fluid_nodes, fluid_nodes + 4 * n);
  m.if_sn.insert(m.if_sn.end(), struct_nodes, struct_nodes + 4 * n);
  ELP_CATCH
}

int elp_model_add_point_source(elp_model *mm, int64_t node_id, double Q) {
  ELP_TRY
  M(mm).sources.push_back({node_id, Q});
  ELP_CATCH
}

int elp_model_add_point_force(elp_model *mm, int64_t node_id, double fx, double fy, double fz) {
  ELP_TRY
  M(mm).forces.push_back({node_id, {fx, fy, fz}});
  ELP_CATCH
}

int elp_model_add_dirichlet(elp_model *mm, int64_t node_id, int dof, double re, double im) {
  ELP_TRY
  M(mm).dirichlet.push_back({node_id, dof, {re, im}});
  ELP_CATCH
}

int elp_model_set_rayleigh(elp_model *mm, double alpha, double beta) {
  ELP_TRY
  auto &m = M(mm);
  ELP_REQUIRE(!m.finalized, ELP_ESTATE, "model already finalized");
  m.ray_a = alpha;
  m.ray_b = beta;
  m.has_rayleigh = true;
  ELP_CATCH
}

int elp_model_set_box(elp_model *mm, const double *origin, const double *length, const int32_t *nel, int order,
                      int robin_sides) {
  ELP_TRY
  auto &m = M(mm);
  ELP_ARG(origin && length && nel && (order == 1 || order == 2), "invalid box description");
  for (int d = 0; d < 3; d++) {
    ELP_ARG(nel[d] > 0 && length[d] > 0, "invalid box description");
    m.box.origin[d] = origin[d];
    m.box.length[d] = length[d];
    m.box.nel[d] = nel[d];
  }
  m.box.order = order;
  m.box.robin_sides = robin_sides;
  m.box.set = true;
  ELP_CATCH
}

int elp_model_get_box(const elp_model *mm, int *set, int *detected, double *origin, double *length, int32_t *nel,
                      int *order, int *robin_sides) {
  ELP_TRY
  ELP_ARG(mm && mm->m, "null model");
  const auto &m = *mm->m;
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  if (set) *set = m.box.set ? 1 : 0;
  if (detected) *detected = m.box.detected ? 1 : 0;
  for (int d = 0; d < 3; d++) {
    if (origin) origin[d] = m.box.set ? m.box.origin[d] : 0.0;
    if (length) length[d] = m.box.set ? m.box.length[d] : 0.0;
    if (nel) nel[d] = m.box.set ? m.box.nel[d] : 0;
  }
  if (order) *order = m.box.set ? m.box.order : 0;
  if (robin_sides) *robin_sides = m.box.set ? m.box.robin_sides : 0;
  ELP_CATCH
}

int elp_model_finalize(elp_model *mm) {
  ELP_TRY
  auto &m = M(mm);
  ELP_HIP(hipSetDevice(mm->m->ctx->device));
  elp::finalize_model(m);
  ELP_CATCH
}

int elp_model_info(const elp_model *mm, int64_t *n_dof, int64_t *nnz, int64_t *device_bytes) {
  ELP_TRY
  ELP_ARG(mm && mm->m, "null model");
  const auto &m = *mm->m;
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  if (n_dof) *n_dof = m.n;
  if (nnz) *nnz = m.nnz;
  if (device_bytes) *device_bytes = elp::model_device_bytes_estimate(m);
  ELP_CATCH
}

int elp_model_dof_map(const elp_model *mm, int64_t *node_ids, int32_t *dofs) {
  ELP_TRY
  ELP_ARG(mm && mm->m && node_ids && dofs, "null argument");
  const auto &m = *mm->m;
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  for (int64_t r = 0; r < m.n_nodes; r++) {
    if (m.node_row[r] < 0) continue;
    int k = 0;
    for (int d = 0; d < 21; d++)
      if (m.node_dofmask[r] & (1u << d)) {
        node_ids[m.node_row[r] + k] = m.sorted_ids[r];
        dofs[m.node_row[r] + k] = d;
        k++;
      }
  }
  ELP_CATCH
}

int elp_assemble(elp_model *mm, double *elapsed_ms) {
  ELP_TRY
  auto &m = M(mm);
  ELP_HIP(hipSetDevice(mm->m->ctx->device));
  auto t0 = std::chrono::steady_clock::now();
  elp::assemble_model(m, m.ctx->stream);
  auto t1 = std::chrono::steady_clock::now();
  if (elapsed_ms) *elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  ELP_CATCH
}

int elp_assemble_timed(elp_model *mm, int repeats, double *kernel_ms) {
  ELP_TRY
  auto &m = M(mm);
  ELP_ARG(repeats > 0, "repeats must be positive");
  ELP_HIP(hipSetDevice(mm->m->ctx->device));
  float ms = elp::time_volume_assembly(m, repeats, m.ctx->stream);
  if (kernel_ms) *kernel_ms = ms;
  ELP_CATCH
}

int elp_get_pattern(const elp_model *mm, int32_t *rowptr, int32_t *colind) {
  ELP_TRY
  ELP_ARG(mm && mm->m, "null model");
  const auto &m = *mm->m;
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  if (rowptr) std::memcpy(rowptr, m.rowptr.data(), (m.n + 1) * sizeof(int32_t));
  if (colind) std::memcpy(colind, m.colind.data(), m.nnz * sizeof(int32_t));
  ELP_CATCH
}

int elp_get_values(const elp_model *mm, int which, double *vals) {
  ELP_TRY
  ELP_ARG(mm && mm->m && vals, "null argument");
  auto &m = *mm->m;
  ELP_REQUIRE(m.assembled, ELP_ESTATE, "model not assembled");
  ELP_HIP(hipSetDevice(mm->m->ctx->device));
  switch (which) {
    case ELP_MAT_K: ELP_HIP(hipMemcpy(vals, m.d_K.p, m.nnz * 8, hipMemcpyDeviceToHost)); break;
    case ELP_MAT_M: ELP_HIP(hipMemcpy(vals, m.d_M.p, m.nnz * 8, hipMemcpyDeviceToHost)); break;
    case ELP_MAT_C:
      ELP_REQUIRE(m.has_rayleigh, ELP_EINVAL, "model has no Rayleigh damping");
      ELP_HIP(hipMemcpy(vals, m.d_C.p, m.nnz * 8, hipMemcpyDeviceToHost));
      break;
    case ELP_MAT_G:
    case ELP_MAT_S: {
      std::memset(vals, 0, m.nnz * 8);
      std::vector<double> xv;
      m.d_x_val.download(xv);
      const uint8_t t = which == ELP_MAT_G ? 0 : 1;
      for (size_t k = 0; k < xv.size(); k++)
        if (m.x_type[k] == t) vals[m.x_fullpos[k]] = xv[k];
      break;
    }
    default: ELP_ARG(false, "unknown matrix");
  }
  ELP_CATCH
}

int elp_get_device_csr(const elp_model *mm, int which, const int32_t **rowptr, const int32_t **colind,
                       const double **vals) {
  ELP_TRY
  ELP_ARG(mm && mm->m, "null model");
  auto &m = *mm->m;
  ELP_REQUIRE(m.assembled, ELP_ESTATE, "model not assembled");
  ELP_ARG(which >= ELP_MAT_K && which <= ELP_MAT_C, "device CSR is available for K, M and C");
  ELP_REQUIRE(which != ELP_MAT_C || m.has_rayleigh, ELP_EINVAL, "model has no Rayleigh damping");
  if (rowptr) *rowptr = m.d_rowptr.p;
  if (colind) *colind = m.d_colind.p;
  if (vals) *vals = which == ELP_MAT_K ? m.d_K.p : (which == ELP_MAT_M ? m.d_M.p : m.d_C.p);
  ELP_CATCH
}

int elp_dynamic_matrix(elp_model *mm, double omega, double z_re, double z_im, double *a, double *f) {
  ELP_TRY
  auto &m = M(mm);
  ELP_HIP(hipSetDevice(mm->m->ctx->device));
  std::complex<double> Z(z_re, z_im);
  if (z_re == 0 && z_im == 0) Z = m.Z;
  elp::dynamic_matrix(m, omega, Z, a, f);
  ELP_CATCH
}

int elp_results_sizes(const elp_model *mm, int64_t *n_nodes, int64_t *n_brick_elements) {
  ELP_TRY
  ELP_ARG(mm && mm->m && n_nodes && n_brick_elements, "null argument");
  const auto &m = *mm->m;
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  *n_nodes = m.n_nodes;
  *n_brick_elements = elp::brick_element_count(m);
  ELP_CATCH
}

int elp_node_results(elp_model *mm, int64_t n_freq, const double *freqs_hz, const double *X, double *cells,
                     double *nodes, double *kernel_ms) {
  ELP_TRY
  auto &m = M(mm);
  ELP_ARG(n_freq > 0 && freqs_hz && X && (cells || nodes), "invalid arguments");
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  ELP_HIP(hipSetDevice(mm->m->ctx->device));
  std::lock_guard<std::mutex> lk(m.mu);
  elp::node_results(m, n_freq, freqs_hz, X, cells, nodes, kernel_ms);
  ELP_CATCH
}

int elp_node_results_noalloc(elp_model *mm, int64_t n_freq, const double *freqs_hz, const double *X_dev,
                             double *nodes_dev) {
  ELP_TRY
  auto &m = M(mm);
  ELP_ARG(n_freq > 0 && freqs_hz && X_dev && nodes_dev, "invalid arguments");
  ELP_REQUIRE(m.finalized, ELP_ESTATE, "model not finalized");
  ELP_HIP(hipSetDevice(mm->m->ctx->device));
  std::lock_guard<std::mutex> lk(m.mu);
  elp::node_results_noalloc(m, n_freq, freqs_hz, X_dev, nodes_dev);
  ELP_CATCH
}

int elp_element_matrices(elp_ctx *ctx, int kind, int64_t n, const double *xyz, double p0, double p1, double p2,
                         double *Ke, double *Me) {
  ELP_TRY
  ELP_ARG(ctx && (n == 0 || (xyz && Ke)), "null argument");
  ELP_HIP(hipSetDevice(ctx->c->device));
  elp::element_matrices_batch(ctx->c.get(), kind, n, xyz, p0, p1, p2, Ke, Me);
  ELP_CATCH
}

int elp_sweep(elp_model *mm, int64_t n_freq, const double *freqs_hz, double z_re, double z_im,
              const elp_solve_opts *opts, double *X, double *relres, int32_t *iters, elp_sweep_stats *stats) {
  ELP_TRY
  auto &m = M(mm);
  ELP_ARG(n_freq > 0 && freqs_hz, "no frequencies");
  ELP_HIP(hipSetDevice(mm->m->ctx->device));
  elp_solve_opts o;
  if (opts) o = *opts;
  else elp::default_opts(o);
  std::complex<double> Z(z_re, z_im);
  if (z_re == 0 && z_im == 0) Z = m.Z;
  std::lock_guard<std::mutex> lk(m.mu);
  if (!m.assembled) elp::assemble_model(m, m.ctx->stream);
  const int64_t bad = elp::sweep(m, n_freq, freqs_hz, Z, o, X, relres, iters, stats, m.ctx->stream);
  if (bad) {
    elp::set_last_error(std::to_string(bad) + " of " + std::to_string(n_freq) +
                        " frequencies did not reach the relative residual " + std::to_string(o.rtol) +
                        " (their best solutions are returned)");
    return ELP_ENOCONV;
  }
  ELP_CATCH
}

}  // extern "C"
