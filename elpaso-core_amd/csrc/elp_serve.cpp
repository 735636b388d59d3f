// Serving loop: continuous batching of frequency-sweep requests on one GPU.
//
// Requests queue in arrival order.  A scheduler thread admits the head of the
// queue whenever a slot is free and the request's device footprint (assembled
// matrices + solver work space for its frequency block) fits in the HBM
// budget next to the requests already running; each admitted request runs on
// its own slot (worker thread + HIP stream), so requests that arrive while
// others are on the GPU overlap with them on the device.  Matrices stay
// resident for the life of the request.  A request's arithmetic does not
// depend on what else runs (private buffers, deterministic kernels), so its
// result is the same alone and in the mix.
#include "elp_serve.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "elp_solver.h"


namespace elp {

struct Job {
  int64_t id = 0;
  elp_model *model = nullptr;
  std::vector<double> freqs;
  std::complex<double> Z;
  elp_solve_opts opts;
  int64_t bytes = 0;
  std::vector<double> X, relres;
  std::vector<int32_t> iters;
  int status = 0;
  std::string error;
  bool done = false;
  std::chrono::steady_clock::time_point t_submit, t_done;
};

struct Server {
  std::shared_ptr<Ctx> ctx_ref;
  Ctx *ctx = nullptr;
  int64_t budget = 0;
  int max_slots = 1;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done;
  std::deque<std::shared_ptr<Job>> pending;
  std::unordered_map<int64_t, std::shared_ptr<Job>> jobs;
  int64_t next_id = 1, in_use = 0, admitted = 0;
  int running = 0;
  int waiters = 0;  // callers inside elp_server_wait
  bool stop = false;
  std::thread scheduler;
  std::vector<std::thread> workers;
  std::deque<std::shared_ptr<Job>> ready;  // admitted, waiting for a slot thread
  std::condition_variable cv_ready;

  void run_job(Job &j, hipStream_t s) {
    Model &m = *j.model->m;
    try {
      ELP_REQUIRE(s, ELP_EHIP, "the server's worker could not create its HIP stream");
      ELP_HIP(hipSetDevice(ctx->device));
      std::lock_guard<std::mutex> lk(m.mu);
      if (!m.assembled) assemble_model(m, s);
      const int64_t nf = (int64_t)j.freqs.size();
      j.X.assign(2 * nf * m.n, 0.0);
      j.relres.assign(nf, 0.0);
      j.iters.assign(nf, 0);
      elp_sweep_stats st{};
      const int64_t bad = sweep(m, nf, j.freqs.data(), j.Z, j.opts, j.X.data(), j.relres.data(), j.iters.data(), &st, s);
      j.status = bad ? ELP_ENOCONV : 0;
      if (bad) j.error = std::to_string(bad) + " of " + std::to_string(nf) + " frequencies did not reach the tolerance";
    } catch (const Error &e) {
      j.status = e.code;
      j.error = e.what();
    } catch (const std::exception &e) {
      j.status = ELP_EINVAL;
      j.error = e.what();
    }
    // the request's device memory is released with its model
    {
      std::lock_guard<std::mutex> lk(m.mu);
      m.solver.reset();
    }
    delete j.model;
    j.model = nullptr;
  }

  void worker_loop() {
    hipStream_t s = nullptr;
    (void)hipSetDevice(ctx->device);
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;  // jobs then fail
    while (true) {
      std::shared_ptr<Job> j;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_ready.wait(lk, [&] { return stop || !ready.empty(); });
        if (ready.empty()) break;
        j = ready.front();
        ready.pop_front();
      }
      run_job(*j, s);
      {
        std::lock_guard<std::mutex> lk(mu);
        j->t_done = std::chrono::steady_clock::now();
        j->done = true;
        in_use -= j->bytes;
        running--;
      }
      cv_done.notify_all();
      cv_sched.notify_all();
    }
    if (s) (void)hipStreamDestroy(s);
  }

  void scheduler_loop() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      cv_sched.wait(lk, [&] {
        return stop || (!pending.empty() && running < max_slots &&
                        (in_use + pending.front()->bytes <= budget || running == 0));
      });
      if (stop) break;
      auto j = pending.front();
      pending.pop_front();
      in_use += j->bytes;
      running++;
      admitted++;
      ready.push_back(j);
      cv_ready.notify_one();
    }
  }
};

}  // namespace elp

struct elp_server {
  elp::Server s;
};

using elp::Error;

#define ELP_TRY try {
#define ELP_CATCH                          \
  }                                        \
  catch (const elp::Error &e) {            \
    elp::set_last_error(e.what());         \
    return e.code;                         \
  }                                        \
  catch (const std::exception &e) {        \
    elp::set_last_error(e.what());         \
    return ELP_EINVAL;                     \
  }                                        \
  return ELP_OK;

extern "C" {

int elp_server_create(elp_ctx *ctx, int64_t hbm_budget_bytes, int max_slots, elp_server **out) {
  ELP_TRY
  ELP_REQUIRE(ctx && out, ELP_EINVAL, "null argument");
  ELP_REQUIRE(hbm_budget_bytes > 0 && max_slots > 0 && max_slots <= 16, ELP_EINVAL, "invalid budget or slot count");
  auto *sv = new elp_server();
  sv->s.ctx = ctx->c.get();
  sv->s.ctx_ref = ctx->c;
  sv->s.budget = hbm_budget_bytes;
  sv->s.max_slots = max_slots;
  for (int i = 0; i < max_slots; i++) sv->s.workers.emplace_back([sv] { sv->s.worker_loop(); });
  sv->s.scheduler = std::thread([sv] { sv->s.scheduler_loop(); });
  *out = sv;
  ELP_CATCH
}

void elp_server_destroy(elp_server *sv) {
  if (!sv) return;
  {
    std::unique_lock<std::mutex> lk(sv->s.mu);
    // drop what never started (its waiters get ELP_ESTATE), let admitted
    // work finish, and wait until no caller is inside elp_server_wait
    sv->s.stop = true;
    for (auto &j : sv->s.pending) {
      delete j->model;
      j->model = nullptr;
      j->status = ELP_ESTATE;
      j->error = "server stopped before the request started";
      j->t_done = std::chrono::steady_clock::now();
      j->done = true;
    }
    sv->s.pending.clear();
    sv->s.cv_done.notify_all();
    sv->s.cv_done.wait(lk, [&] { return sv->s.running == 0 && sv->s.waiters == 0; });
  }
  sv->s.cv_sched.notify_all();
  sv->s.cv_ready.notify_all();
  sv->s.scheduler.join();
  for (auto &w : sv->s.workers) w.join();
  delete sv;
}

int elp_server_submit(elp_server *sv, elp_model *m, int64_t n_freq, const double *freqs_hz, double z_re,
                      double z_im, const elp_solve_opts *opts, int64_t *job_id) {
  ELP_TRY
  ELP_REQUIRE(sv && m && m->m && freqs_hz && n_freq > 0 && job_id, ELP_EINVAL, "invalid argument");
  ELP_REQUIRE(m->m->finalized, ELP_ESTATE, "submit a finalized model");
  auto j = std::make_shared<elp::Job>();
  j->model = m;
  j->freqs.assign(freqs_hz, freqs_hz + n_freq);
  for (double f : j->freqs) ELP_REQUIRE(f > 0, ELP_EINVAL, "frequencies must be positive");
  j->Z = (z_re == 0 && z_im == 0) ? m->m->Z : std::complex<double>(z_re, z_im);
  if (opts) j->opts = *opts;
  else elp::default_opts(j->opts);
  const int B = j->opts.precision == ELP_PREC_FP64 ? 1 : j->opts.block;
  j->bytes = elp::model_device_bytes_estimate(*m->m) + elp::solver_bytes_estimate(*m->m, B, j->opts.precision);
  j->t_submit = std::chrono::steady_clock::now();
  {
    std::lock_guard<std::mutex> lk(sv->s.mu);
    ELP_REQUIRE(!sv->s.stop, ELP_ESTATE, "server stopped");
    j->id = sv->s.next_id++;
    sv->s.jobs[j->id] = j;
    sv->s.pending.push_back(j);
    *job_id = j->id;
  }
  sv->s.cv_sched.notify_all();
  ELP_CATCH
}

int elp_server_poll(elp_server *sv, int64_t job_id) {
  if (!sv) return ELP_EINVAL;
  std::lock_guard<std::mutex> lk(sv->s.mu);
  auto it = sv->s.jobs.find(job_id);
  if (it == sv->s.jobs.end()) return ELP_EINVAL;
  return it->second->done ? 1 : 0;
}

int elp_server_wait(elp_server *sv, int64_t job_id, double *X, double *relres, int32_t *iters, double *latency_s) {
  ELP_TRY
  ELP_REQUIRE(sv, ELP_EINVAL, "null server");
  std::shared_ptr<elp::Job> j;
  {
    std::unique_lock<std::mutex> lk(sv->s.mu);
    auto it = sv->s.jobs.find(job_id);
    ELP_REQUIRE(it != sv->s.jobs.end(), ELP_EINVAL, "unknown job");
    j = it->second;
    sv->s.waiters++;
    sv->s.cv_done.wait(lk, [&] { return j->done; });
    sv->s.waiters--;
    sv->s.jobs.erase(job_id);
  }
  sv->s.cv_done.notify_all();  // a destroy may be waiting for the waiters
  if (latency_s) *latency_s = std::chrono::duration<double>(j->t_done - j->t_submit).count();
  if (j->status != 0 && j->status != ELP_ENOCONV) throw Error(j->status, j->error);
  if (X) std::memcpy(X, j->X.data(), j->X.size() * sizeof(double));
  if (relres) std::memcpy(relres, j->relres.data(), j->relres.size() * sizeof(double));
  if (iters) std::memcpy(iters, j->iters.data(), j->iters.size() * sizeof(int32_t));
  if (j->status == ELP_ENOCONV) throw Error(j->status, j->error);  // results are filled
  ELP_CATCH
}

int elp_server_stats(elp_server *sv, int64_t *admitted, int64_t *queued, int64_t *bytes_in_use) {
  ELP_TRY
  ELP_REQUIRE(sv, ELP_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(sv->s.mu);
  if (admitted) *admitted = sv->s.admitted;
  if (queued) *queued = (int64_t)sv->s.pending.size();
  if (bytes_in_use) *bytes_in_use = sv->s.in_use;
  ELP_CATCH
}

}  // extern "C"
