// Serving loop: continuous batching of frequency-sweep requests on one GPU.
//
// Requests queue in arrival order.  A scheduler admits the head of the queue
// whenever fewer than max_slots requests are resident and its device
// footprint (assembled matrices + solver work space) fits in the HBM budget
// next to the resident ones.  An admitted request is set up (assembly if
// needed, solver state, work space) on a set-up thread with its own stream,
// so a large request's set-up does not stall the others, and then joins the
// group of resident requests of its precision and slot count.  A group's
// driver thread advances all of its requests together: each step is one
// COCG iteration of every request, with the operator, the dot products and
// the vector updates of all of them fused into single launches (the frequency
// slots of different requests share each launch), and each request's slots
// refilled with its next frequencies as they converge.  Requests that arrive
// while others run join the group at the next step.  A request's arithmetic
// is the same in any mix (each request keeps its own partition of every
// launch), so its result is the same alone and in the mix.  Requests of the
// plain FP64 precision (one frequency at a time) run alone on the set-up
// thread.
#include "elp_serve.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
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
  int group = -1;  // key of the group it runs in
  std::chrono::steady_clock::time_point t_submit, t_done;
};

struct Server;

// one group: a driver thread with its own stream
struct Group {
  Server *sv = nullptr;
  int key = 0;
  hipStream_t s = nullptr;
  std::unique_ptr<GroupDriver> drv;
  std::deque<std::unique_ptr<GroupItem>> inbox;
  std::thread th;
  void loop();
};

struct Server {
  std::shared_ptr<Ctx> ctx_ref;
  Ctx *ctx = nullptr;
  int64_t budget = 0;
  int max_slots = 1;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_setup, cv_group;
  std::deque<std::shared_ptr<Job>> pending;   // submitted, not admitted
  std::deque<std::shared_ptr<Job>> to_setup;  // admitted, waiting for set-up
  std::unordered_map<int64_t, std::shared_ptr<Job>> jobs;
  std::map<int, std::unique_ptr<Group>> groups;
  int64_t next_id = 1, in_use = 0, admitted = 0;
  int running = 0;     // admitted and not finished
  int setting_up = 0;  // taken by the set-up thread, not yet in a group or ended
  int waiters = 0;     // callers inside elp_server_wait
  bool stop = false;
  std::thread scheduler, setup;

  // a request ended (under mu): results are in the job
  void complete(Job &j, int status, const std::string &err) {
    if (j.done) return;
    j.status = status;
    j.error = err;
    if (j.model) {
      delete j.model;  // the request's device memory goes with its model
      j.model = nullptr;
    }
    j.t_done = std::chrono::steady_clock::now();
    j.done = true;
    in_use -= j.bytes;
    running--;
    cv_done.notify_all();
    cv_sched.notify_all();
  }

  // (under mu) a request left the set-up thread; groups waiting to shut
  // down look again
  void setup_done() {
    setting_up--;
    cv_group.notify_all();
  }
  // (under mu) no request can reach a group any more
  bool setup_drained() const { return stop && to_setup.empty() && setting_up == 0; }

  static int unconverged(const Job &j) {
    int bad = 0;
    for (double r : j.relres) bad += !(r <= j.opts.rtol);
    return bad;
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
      to_setup.push_back(j);
      cv_setup.notify_one();
    }
  }

  void setup_loop() {
    hipStream_t s = nullptr;
    (void)hipSetDevice(ctx->device);
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    while (true) {
      std::shared_ptr<Job> j;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_setup.wait(lk, [&] { return stop || !to_setup.empty(); });
        if (to_setup.empty()) break;
        j = to_setup.front();
        to_setup.pop_front();
        setting_up++;
      }
      Model &m = *j->model->m;
      const int64_t nf = (int64_t)j->freqs.size();
      std::unique_ptr<GroupItem> item;
      int key = -1;
      try {
        ELP_REQUIRE(s, ELP_EHIP, "the server could not create its set-up stream");
        ELP_HIP(hipSetDevice(ctx->device));
        std::lock_guard<std::mutex> mlk(m.mu);
        if (!m.assembled) assemble_model(m, s);
        j->X.assign(2 * nf * m.n, 0.0);
        j->relres.assign(nf, 0.0);
        j->iters.assign(nf, 0);
        item = group_prepare(m, nf, j->freqs.data(), j->Z, j->opts, j->X.data(), j->relres.data(),
                             j->iters.data(), s, &key);
        if (!item) {
          // one frequency at a time: solved here, alone
          elp_sweep_stats st{};
          sweep(m, nf, j->freqs.data(), j->Z, j->opts, j->X.data(), j->relres.data(), j->iters.data(), &st, s);
        }
      } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(mu);
        complete(*j, e.code, e.what());
        setup_done();
        continue;
      } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(mu);
        complete(*j, ELP_EINVAL, e.what());
        setup_done();
        continue;
      }
      std::lock_guard<std::mutex> lk(mu);
      setup_done();
      if (!item) {
        const int bad = unconverged(*j);
        complete(*j, bad ? ELP_ENOCONV : 0,
                 bad ? std::to_string(bad) + " frequencies did not reach the tolerance" : std::string());
        continue;
      }
      item->tag = j->id;
      j->group = key;
      auto &g = groups[key];
      if (!g) {
        g.reset(new Group());
        g->sv = this;
        g->key = key;
        g->th = std::thread([gp = g.get()] { gp->loop(); });
      }
      g->inbox.push_back(std::move(item));
      cv_group.notify_all();
    }
    if (s) (void)hipStreamDestroy(s);
  }
};

void Group::loop() {
  Server &S = *sv;
  (void)hipSetDevice(S.ctx->device);
  std::string init_err;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) init_err = "cannot create a stream";
  if (init_err.empty()) {
    try {
      drv = group_driver(key, S.ctx, s);
    } catch (const std::exception &e) {
      init_err = e.what();
    }
  }
  std::vector<std::unique_ptr<GroupItem>> fresh, finished;
  while (true) {
    {
      std::unique_lock<std::mutex> lk(S.mu);
      // on shutdown a group ends only once the set-up thread can hand it no
      // more requests (an admitted request may still be in set-up)
      S.cv_group.wait(lk, [&] { return S.setup_drained() || !inbox.empty() || (drv && !drv->empty()); });
      if (S.setup_drained() && inbox.empty() && (!drv || drv->empty())) break;
      while (!inbox.empty()) {
        fresh.push_back(std::move(inbox.front()));
        inbox.pop_front();
      }
    }
    finished.clear();
    std::string err;
    int code = 0;
    std::vector<int64_t> failed;
    try {
      ELP_REQUIRE(init_err.empty(), ELP_EHIP, init_err);
      for (auto &it : fresh) {
        const int64_t tag = it->tag;
        try {
          drv->add(std::move(it));
        } catch (const Error &e) {
          failed.push_back(tag);
          code = e.code;
          err = e.what();
        }
      }
      fresh.clear();
      drv->step(finished);
    } catch (const Error &e) {
      code = e.code;
      err = e.what();
    } catch (const std::exception &e) {
      code = ELP_EINVAL;
      err = e.what();
    }
    std::lock_guard<std::mutex> lk(S.mu);
    for (auto &it : finished) {
      auto jt = S.jobs.find(it->tag);
      it.reset();  // the request's solver memory is released before its model
      if (jt == S.jobs.end()) continue;
      const int bad = Server::unconverged(*jt->second);
      S.complete(*jt->second, bad ? ELP_ENOCONV : 0,
                 bad ? std::to_string(bad) + " frequencies did not reach the tolerance" : std::string());
    }
    for (int64_t tag : failed) {
      auto jt = S.jobs.find(tag);
      if (jt != S.jobs.end()) S.complete(*jt->second, code, err);
    }
    if (code && failed.empty()) {
      // a failure of the group's step: every request still in the group
      // ends with the error
      drv.reset();
      fresh.clear();
      for (auto &kv : S.jobs)
        if (!kv.second->done && kv.second->group == key) S.complete(*kv.second, code, err);
      try {
        drv = group_driver(key, S.ctx, s);
      } catch (const std::exception &e) {
        init_err = e.what();
      }
    }
  }
  drv.reset();
  if (s) (void)hipStreamDestroy(s);
}

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
  ELP_REQUIRE(hbm_budget_bytes > 0 && max_slots > 0 && max_slots <= 64, ELP_EINVAL, "invalid budget or slot count");
  auto *sv = new elp_server();
  sv->s.ctx = ctx->c.get();
  sv->s.ctx_ref = ctx->c;
  sv->s.budget = hbm_budget_bytes;
  sv->s.max_slots = max_slots;
  sv->s.setup = std::thread([sv] { sv->s.setup_loop(); });
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
    sv->s.cv_setup.notify_all();
    sv->s.cv_group.notify_all();
    sv->s.cv_done.wait(lk, [&] { return sv->s.running == 0 && sv->s.waiters == 0; });
  }
  sv->s.cv_sched.notify_all();
  sv->s.cv_setup.notify_all();
  sv->s.cv_group.notify_all();
  sv->s.scheduler.join();
  sv->s.setup.join();
  for (auto &kv : sv->s.groups) kv.second->th.join();
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
    for (auto &kv : sv->s.jobs)  // a model belongs to one request at a time
      ELP_REQUIRE(kv.second->model != m, ELP_EINVAL, "model already submitted");
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
