// Serving runtime: a request loop that batches BEM solves as they arrive on one GPU.
//
// Requests (geometry, refinement level, degree, operator, right-hand sides, solver settings) are
// queued by bmi_server_submit and admitted by a scheduler thread (smallest problems first, any
// request older than a few seconds before all others) while
//   - fewer than `max_active` batches are on the GPU, and
//   - the device-memory budget holds the batch's workspace plus, when the operator is not yet
//     resident, its assembled H2 state (estimated from the block tree before anything is
//     allocated); idle resident operators are evicted least recently used first.
// Queued requests with the same operator and solver settings are merged into one batched solve
// (up to `max_batch_rhs` right-hand sides).  Admitted batches hold their memory and are advanced in
// time slices by one device thread: a GMRES batch runs `slice_iterations` steps (whole restart
// cycles; a restarted GMRES resumed from its iterate continues exactly as it would have), then the
// device turns to the next admitted batch, so a short solve that arrives while a long one runs
// starts after at most one slice of each batch ahead of it.  CG batches run to the end in one
// slice (a restarted CG converges more slowly).  Every HIP call of the server is made by the device
// thread; the scheduler thread does host work only (geometry, spaces, memory estimates), and
// operators it evicts are released by the device thread.  Assembled operators stay resident
// between requests (keyed by geometry file, scale, level, degree and operator), so a repeated
// geometry skips assembly.
//
// There is no counterpart in the reference (Bembel solves one problem per program); the
// operator, right-hand side and solver calls are the library's own C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "bmi_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Operator {
  std::string key;
  bmi_geometry* g = nullptr;
  bmi_space* s = nullptr;
  bmi_h2* h = nullptr;
  bmi_op_desc op{};
  int64_t bytes = 0;  // resident device bytes (estimate)
  int64_t ndof = 0, nsuper = 0;
  int users = 0;  // admitted batches using the operator (not evicted while > 0)
  double last_used = 0.0;
  Operator() = default;
  Operator(const Operator&) = delete;
  Operator& operator=(const Operator&) = delete;
  ~Operator() {
    if (h) bmi_h2_destroy(h);
    if (s) bmi_space_destroy(s);
    if (g) bmi_geometry_destroy(g);
  }
};

struct Request {
  int64_t id = 0;
  std::string path;
  double scale[3] = {1, 1, 1};
  int M = 0, P = 0, form = 0;
  bmi_op_desc op{};
  int data_fn = 0, nrhs = 1;
  double data_kappa = 0.0;
  std::vector<double> params;
  double tol = 1e-10;
  int restart = 30, maxit = 1000;
  std::string key, batch_key;
  // bookkeeping
  std::shared_ptr<Operator> pending;  // operator built by the scheduler, not yet resident
  int64_t op_bytes = -1, ws_bytes = 0;
  int state = 0;  // 0 queued, 1 running, 2 finished
  bmi_status status = BMI_OK;
  std::string error;
  double t_submit = 0, t_start = -1, t_end = -1;
  std::vector<double> x;
  std::vector<int> iters;
  std::vector<double> resid;
  int64_t ndof = 0;
  int batch_requests = 0, batch_rhs = 0, cache_hit = 0, slices = 0;
};

// an admitted batch: its device buffers and progress between time slices
struct Job {
  std::vector<std::shared_ptr<Request>> batch;
  std::shared_ptr<Operator> op;
  int64_t ws = 0;
  hipStream_t stream = nullptr;
  double* dB = nullptr;
  double* dX = nullptr;
  bmi_gmres_task* task = nullptr;  // GMRES in flight on `stream`
  int rhs = 0, S = 1, slices = 0, cache_hit = 0;
  bool started = false;
  bmi_status st = BMI_OK;
  std::string err;
  std::vector<int> iters;
  std::vector<double> resid;
  std::vector<double> hx;
};

std::string operator_key(const Request& r) {
  std::ostringstream k;
  k.precision(17);
  k << r.path << '|' << r.scale[0] << ',' << r.scale[1] << ',' << r.scale[2] << '|' << r.M << ',' << r.P << ','
    << r.form << '|' << r.op.kind << ',' << r.op.kappa_re << ',' << r.op.kappa_im << ',' << r.op.n_cheb << ','
    << r.op.eta << ',' << r.op.min_cluster_level << ',' << r.op.precision;
  return k.str();
}

}  // namespace

struct bmi_server {
  bmi_server_config cfg{};
  Clock::time_point t0 = Clock::now();
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<int64_t, std::shared_ptr<Request>> reqs;
  std::map<std::string, std::shared_ptr<Operator>> ops;
  std::deque<std::shared_ptr<Job>> ready;              // admitted batches, in turn order
  std::vector<std::shared_ptr<Operator>> graveyard;    // evicted, released by the device thread
  int64_t used_bytes = 0;
  int active = 0;
  int64_t next_id = 1;
  bool stop = false;
  std::thread scheduler, device;
  int64_t submitted = 0, finished = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  void run_scheduler();
  void run_device();
  // device thread, without the lock: 0 waiting, 1 finished, 2 work enqueued
  int step(Job& j, const std::vector<std::shared_ptr<Job>>& live);
  void release(Job& j);                   // device thread: solution to the host, buffers freed
  void finish(const std::shared_ptr<Job>& j);  // with the lock
  bool estimate(Request& r);              // called without the lock
};

// geometry, space and memory estimate of a queued request whose operator is not resident
bool bmi_server::estimate(Request& r) {
  auto op = std::make_shared<Operator>();
  op->key = r.key;
  op->op = r.op;
  const bool scaled = r.scale[0] != 1.0 || r.scale[1] != 1.0 || r.scale[2] != 1.0;
  bmi_status st = bmi_geometry_load(r.path.c_str(), scaled ? r.scale : nullptr, &op->g);
  if (st == BMI_OK) st = bmi_space_create(op->g, r.M, r.P, r.form, &op->s);
  int64_t b = 0;
  if (st == BMI_OK) st = bmi_h2_estimate_bytes(op->s, &r.op, 0, &b);
  if (st != BMI_OK) {
    r.status = st;
    r.error = bmi_last_error();
    return false;
  }
  int64_t info[8];
  bmi_space_info(op->s, info);
  op->ndof = info[5];
  op->nsuper = info[4];
  op->bytes = b;
  r.op_bytes = b;
  r.pending = op;
  return true;
}

void bmi_server::run_scheduler() {
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    cv.wait(lk, [&] { return stop || (!queue.empty() && active < cfg.max_active); });
    if (stop) return;
    // estimates of queued requests without a resident operator (host work, lock released)
    for (size_t qi = 0; qi < queue.size(); ++qi) {
      std::shared_ptr<Request> r = queue[qi];
      if (r->op_bytes >= 0 || ops.count(r->key)) continue;
      lk.unlock();
      const bool ok = estimate(*r);
      lk.lock();
      if (!ok) {
        r->state = 2;
        r->t_end = now();
        ++finished;
        queue.erase(std::find(queue.begin(), queue.end(), r));
        cv.notify_all();
        qi = static_cast<size_t>(-1);  // restart the scan: the queue changed
        continue;
      }
      if (r->op_bytes + r->ws_bytes > cfg.hbm_budget_bytes) {
        r->status = BMI_ADMISSION_REJECTED;
        r->error = "request needs more device memory than the server budget";
        r->state = 2;
        r->t_end = now();
        ++finished;
        queue.erase(std::find(queue.begin(), queue.end(), r));
        cv.notify_all();
        qi = static_cast<size_t>(-1);
      }
    }
    // admission order: requests that waited longer than aging_s first (in arrival order), then
    // the smallest problems first (degrees of freedom x right-hand sides), so that short solves
    // do not queue behind long ones
    {
      const double tnow = now(), aging_s = 5.0;
      auto size_of = [&](const std::shared_ptr<Request>& q) {
        auto oi = ops.find(q->key);
        const int64_t nd = oi != ops.end() ? oi->second->ndof : (q->pending ? q->pending->ndof : -1);
        return nd < 0 ? std::numeric_limits<double>::max() : static_cast<double>(nd) * q->nrhs;
      };
      std::stable_sort(queue.begin(), queue.end(), [&](const std::shared_ptr<Request>& a, const std::shared_ptr<Request>& b) {
        const bool oa = tnow - a->t_submit > aging_s, ob = tnow - b->t_submit > aging_s;
        if (oa != ob) return oa;
        if (oa) return a->t_submit < b->t_submit;
        return size_of(a) < size_of(b);
      });
    }
    bool admitted = false;
    for (size_t qi = 0; qi < queue.size() && active < cfg.max_active; ++qi) {
      std::shared_ptr<Request> r = queue[qi];
      auto it = ops.find(r->key);
      if (it != ops.end() && it->second->users > 0) continue;  // its operator is in use: wait
      if (it == ops.end() && r->op_bytes < 0) continue;
      // batch: queued requests with the same operator, data kind and solver settings
      std::vector<std::shared_ptr<Request>> batch;
      int rhs = 0;
      for (auto& q : queue)
        if (q->batch_key == r->batch_key && rhs + q->nrhs <= std::max(cfg.max_batch_rhs, r->nrhs)) {
          batch.push_back(q);
          rhs += q->nrhs;
        }
      const Operator& o = it != ops.end() ? *it->second : *r->pending;
      const int S = (r->op.kind == BMI_LAPLACE_SL) ? 1 : 2;
      // solver vectors, plus the product's super-space and multipole workspace per column
      const int64_t ws = static_cast<int64_t>(rhs) * S * 8 *
                         ((r->op.kind == BMI_LAPLACE_SL ? 6 : r->restart + 6) * o.ndof + 5 * o.nsuper);
      int64_t need = ws + (it == ops.end() ? r->op_bytes : 0);
      // evict idle resident operators, least recently used first
      while (used_bytes + need > cfg.hbm_budget_bytes) {
        std::shared_ptr<Operator> victim;
        for (auto& o : ops)
          if (o.second->users == 0 && o.second->key != r->key &&
              (!victim || o.second->last_used < victim->last_used))
            victim = o.second;
        if (!victim) break;
        used_bytes -= victim->bytes;
        ops.erase(victim->key);
        graveyard.push_back(std::move(victim));  // its device memory is released by the device thread
        cv.notify_all();
      }
      if (used_bytes + need > cfg.hbm_budget_bytes) continue;
      std::shared_ptr<Operator> op;
      if (it != ops.end()) {
        op = it->second;
      } else {
        op = r->pending;
        ops[r->key] = op;
        used_bytes += op->bytes;
      }
      ++op->users;
      used_bytes += ws;
      ++active;
      const double t = now();
      for (auto& q : batch) {
        q->state = 1;
        q->t_start = t;
        q->pending.reset();
        queue.erase(std::find(queue.begin(), queue.end(), q));
      }
      qi = static_cast<size_t>(-1);
      admitted = true;
      auto job = std::make_shared<Job>();
      job->batch = std::move(batch);
      job->op = op;
      job->ws = ws;
      ready.push_back(std::move(job));
      cv.notify_all();
    }
    if (!admitted) cv.wait_for(lk, std::chrono::milliseconds(5));
  }
}

void bmi_server::run_device() {
  std::vector<std::shared_ptr<Job>> live;  // admitted batches on the device (this thread only)
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    if (live.empty()) cv.wait(lk, [&] { return stop || !ready.empty() || !graveyard.empty(); });
    if (!graveyard.empty()) {
      std::vector<std::shared_ptr<Operator>> g;
      g.swap(graveyard);
      lk.unlock();
      g.clear();  // bmi_h2_destroy and friends, outside the lock
      lk.lock();
      continue;
    }
    while (!ready.empty()) {
      live.push_back(ready.front());
      ready.pop_front();
    }
    if (live.empty()) {
      if (stop) return;
      continue;
    }
    lk.unlock();
    // one round over the admitted batches: each enqueues its next work if it is ready, so the
    // batches' kernels are in flight on the device together
    bool progress = false;
    std::vector<std::shared_ptr<Job>> done;
    for (auto& j : live) {
      const int r = step(*j, live);
      if (r == 1) done.push_back(j);
      progress = progress || r != 0;
    }
    lk.lock();
    for (auto& j : done) {
      finish(j);
      live.erase(std::find(live.begin(), live.end(), j));
    }
    if (!done.empty()) cv.notify_all();
    // nothing ready on the device: a short wait (new admissions and evictions wake it); short,
    // since a task waiting for its flags leaves the device idle until its next turn
    if (!progress) cv.wait_for(lk, std::chrono::microseconds(20));
  }
}

// one turn of a batch: on the first, assembly (when the operator is not resident), buffers and
// right-hand sides, and the start of its GMRES task; then the task's next work when its
// convergence flags have arrived. CG batches, and every batch with slice_iterations == 0, run
// to the end in their first turn. A batch waits while another admitted batch uses the same
// operator (the product workspace belongs to the operator).
int bmi_server::step(Job& j, const std::vector<std::shared_ptr<Job>>& live) {
  const Request& r0 = *j.batch[0];
  const bool gm = r0.op.kind != BMI_LAPLACE_SL;
  auto fail = [&](bmi_status s, const char* what) {
    if (j.st == BMI_OK) j.st = s, j.err = what ? std::string(what) : std::string(bmi_last_error());
  };
  if (!j.started) {
    for (const auto& k : live)
      if (k.get() != &j && k->started && k->op == j.op) return 0;
    j.started = true;
    j.S = gm ? 2 : 1;
    for (auto& q : j.batch) j.rhs += q->nrhs;
    j.iters.assign(j.rhs, 0);
    j.resid.assign(j.rhs, 0.0);
    j.cache_hit = j.op->h != nullptr ? 1 : 0;
    if (hipStreamCreateWithFlags(&j.stream, hipStreamNonBlocking) != hipSuccess) {
      j.stream = nullptr;
      fail(BMI_DEVICE_ERROR, "stream creation failed");
    }
    if (j.st == BMI_OK && !j.op->h) {
      const bmi_status s = bmi_h2_assemble(j.op->s, &j.op->op, j.stream, &j.op->h);
      if (s != BMI_OK) fail(s, nullptr);
      if (s == BMI_OK) bmi_h2_use_graphs(j.op->h, 0);
    }
    const int64_t n = j.op->ndof;
    if (j.st == BMI_OK && (hipMalloc(&j.dB, sizeof(double) * j.S * n * j.rhs) != hipSuccess ||
                           hipMalloc(&j.dX, sizeof(double) * j.S * n * j.rhs) != hipSuccess))
      fail(BMI_OUT_OF_MEMORY, "workspace allocation failed");
    if (j.st == BMI_OK && bmi::poison_allocations()) {
      const size_t bytes = sizeof(double) * j.S * n * j.rhs;
      if (hipMemsetAsync(j.dB, 0xFF, bytes, nullptr) != hipSuccess ||
          hipMemsetAsync(j.dX, 0xFF, bytes, nullptr) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess)
        fail(BMI_DEVICE_ERROR, "poison fill failed");
    }
    int col = 0;
    for (auto& q : j.batch) {
      if (j.st != BMI_OK) break;
      const bmi_status s = bmi_rhs_assemble(j.op->s, q->data_fn, q->params.data(), q->nrhs, q->data_kappa,
                                            j.dB + static_cast<int64_t>(j.S) * n * col, BMI_MEM_DEVICE, j.stream);
      if (s != BMI_OK) fail(s, nullptr);
      col += q->nrhs;
    }
    if (j.st != BMI_OK) {
      release(j);
      return 1;
    }
    ++j.slices;
    if (!gm) {
      const bmi_status s = bmi_solve_cg(j.op->h, j.rhs, j.dB, j.dX, r0.tol, r0.maxit, 1, j.iters.data(),
                                        j.resid.data(), BMI_MEM_DEVICE, j.stream);
      if (s != BMI_OK) fail(s, s == BMI_NOT_CONVERGED ? "solver did not reach the tolerance" : nullptr);
      release(j);
      return 1;
    }
    const bmi_status s = bmi_gmres_start(j.op->h, j.rhs, j.dB, j.dX, r0.tol, r0.restart, r0.maxit, 1, j.stream,
                                         &j.task);
    if (s != BMI_OK) {
      fail(s, nullptr);
      release(j);
      return 1;
    }
    if (cfg.slice_iterations > 0) return 2;
  }
  int d = 1;
  if (cfg.slice_iterations > 0) {
    const bmi_status s = bmi_gmres_poll(j.task, &d);
    if (s != BMI_OK) fail(s, nullptr), d = 1;
    if (d == 0) return 0;
    if (d == 2) {
      ++j.slices;
      return 2;
    }
  }
  const bmi_status s = bmi_gmres_finish(j.task, j.iters.data(), j.resid.data());
  j.task = nullptr;
  if (s != BMI_OK) fail(s, s == BMI_NOT_CONVERGED ? "solver did not reach the tolerance" : nullptr);
  release(j);
  return 1;
}

void bmi_server::release(Job& j) {
  if (j.task) {
    (void)bmi_gmres_finish(j.task, nullptr, nullptr);
    j.task = nullptr;
  }
  if (j.dX && (j.st == BMI_OK || j.st == BMI_NOT_CONVERGED)) {
    j.hx.resize(static_cast<size_t>(j.S) * j.op->ndof * j.rhs);
    if (hipMemcpyAsync(j.hx.data(), j.dX, j.hx.size() * sizeof(double), hipMemcpyDeviceToHost, j.stream) !=
            hipSuccess ||
        hipStreamSynchronize(j.stream) != hipSuccess) {
      j.st = BMI_DEVICE_ERROR;
      j.err = "copy of the solution failed";
      j.hx.clear();
    }
  }
  if (j.stream) (void)hipStreamSynchronize(j.stream);
  if (j.dB) (void)hipFree(j.dB);
  if (j.dX) (void)hipFree(j.dX);
  if (j.stream) (void)hipStreamDestroy(j.stream);
  j.dB = j.dX = nullptr;
  j.stream = nullptr;
}

void bmi_server::finish(const std::shared_ptr<Job>& jp) {
  Job& j = *jp;
  const double t = now();
  const int64_t n = j.op->ndof;
  int col = 0;
  for (auto& q : j.batch) {
    q->status = j.st;
    q->error = j.err;
    q->ndof = n;
    q->batch_requests = static_cast<int>(j.batch.size());
    q->batch_rhs = j.rhs;
    q->cache_hit = j.cache_hit;
    q->slices = j.slices;
    if (!j.hx.empty())
      q->x.assign(j.hx.begin() + static_cast<int64_t>(j.S) * n * col,
                  j.hx.begin() + static_cast<int64_t>(j.S) * n * (col + q->nrhs));
    if (static_cast<int>(j.iters.size()) >= col + q->nrhs) {
      q->iters.assign(j.iters.begin() + col, j.iters.begin() + col + q->nrhs);
      q->resid.assign(j.resid.begin() + col, j.resid.begin() + col + q->nrhs);
    }
    q->t_end = t;
    q->state = 2;
    col += q->nrhs;
    ++finished;
  }
  j.hx.clear();
  used_bytes -= j.ws;
  --j.op->users;
  j.op->last_used = t;
  if (!j.op->h) {  // assembly failed: forget the operator
    used_bytes -= j.op->bytes;
    ops.erase(j.op->key);
    graveyard.push_back(j.op);
  }
  j.op.reset();
  --active;
}

extern "C" {

bmi_status bmi_server_config_default(bmi_server_config* cfg) {
  if (!cfg) return BMI_INVALID_ARG;
  cfg->hbm_budget_bytes = 200.0 * (1ull << 30);
  cfg->max_active = 4;
  cfg->max_batch_rhs = 64;
  cfg->slice_iterations = 120;
  return BMI_OK;
}

bmi_status bmi_server_create(const bmi_server_config* cfg, bmi_server** out) {
  try {
    BMI_REQUIRE(cfg && out, BMI_INVALID_ARG, "null argument");
    BMI_REQUIRE(cfg->max_active >= 1 && cfg->max_batch_rhs >= 1 && cfg->hbm_budget_bytes > 0 &&
                    cfg->slice_iterations >= 0,
                BMI_INVALID_ARG, "bad server configuration");
    auto s = std::make_unique<bmi_server>();
    s->cfg = *cfg;
    s->scheduler = std::thread(&bmi_server::run_scheduler, s.get());
    s->device = std::thread(&bmi_server::run_device, s.get());
    *out = s.release();
    return BMI_OK;
  } catch (const bmi::Error& e) {
    bmi::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    bmi::set_last_error(e.what());
    return BMI_DEVICE_ERROR;
  }
}

bmi_status bmi_server_submit(bmi_server* srv, const bmi_request* rq, int64_t* id) {
  try {
    BMI_REQUIRE(srv && rq && id && rq->geometry_path, BMI_INVALID_ARG, "null argument");
    BMI_REQUIRE(rq->nrhs >= 1 && rq->data_params, BMI_INVALID_ARG, "no right-hand sides");
    BMI_REQUIRE(rq->op.kind == BMI_LAPLACE_SL || rq->op.kind == BMI_HELMHOLTZ_SL, BMI_UNSUPPORTED,
                "the server solves Laplace and Helmholtz single-layer problems");
    BMI_REQUIRE((rq->op.kind == BMI_LAPLACE_SL) == (rq->data_fn == 0), BMI_INVALID_ARG,
                "Laplace takes point sources (fn 0), Helmholtz plane waves (fn 1)");
    BMI_REQUIRE(rq->tol > 0 && rq->maxit >= 1 && rq->restart >= 1, BMI_INVALID_ARG, "bad solver settings");
    auto r = std::make_shared<Request>();
    r->path = rq->geometry_path;
    for (int k = 0; k < 3; ++k) r->scale[k] = rq->scale[k] != 0.0 ? rq->scale[k] : 1.0;
    r->M = rq->refinement_level;
    r->P = rq->polynomial_degree;
    r->form = BMI_FORM_DISCONTINUOUS;
    r->op = rq->op;
    r->data_fn = rq->data_fn;
    r->nrhs = rq->nrhs;
    r->data_kappa = rq->data_kappa;
    r->params.assign(rq->data_params, rq->data_params + 3 * rq->nrhs);
    r->tol = rq->tol;
    r->restart = rq->restart;
    r->maxit = rq->maxit;
    r->key = operator_key(*r);
    std::ostringstream bk;
    bk.precision(17);
    bk << r->key << '#' << r->data_fn << ',' << r->data_kappa << ',' << r->tol << ',' << r->restart << ',' << r->maxit;
    r->batch_key = bk.str();
    std::lock_guard<std::mutex> lk(srv->mu);
    r->id = srv->next_id++;
    r->t_submit = srv->now();
    srv->reqs[r->id] = r;
    srv->queue.push_back(r);
    ++srv->submitted;
    *id = r->id;
    srv->cv.notify_all();
    return BMI_OK;
  } catch (const bmi::Error& e) {
    bmi::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    bmi::set_last_error(e.what());
    return BMI_DEVICE_ERROR;
  }
}

bmi_status bmi_server_wait(bmi_server* srv, int64_t id, double timeout_s, bmi_result_info* info) {
  if (!srv) return BMI_INVALID_ARG;
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(id);
  if (it == srv->reqs.end()) {
    bmi::set_last_error("unknown request id");
    return BMI_INVALID_ARG;
  }
  std::shared_ptr<Request> r = it->second;
  const auto done = [&] { return r->state == 2; };
  if (timeout_s < 0)
    srv->cv.wait(lk, done);
  else if (!srv->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
    return BMI_PENDING;
  if (info) {
    info->status = r->status;
    info->ndof = r->ndof;
    info->nrhs = r->nrhs;
    info->submit_s = r->t_submit;
    info->start_s = r->t_start;
    info->end_s = r->t_end;
    info->iterations = r->iters.empty() ? 0 : *std::max_element(r->iters.begin(), r->iters.end());
    info->residual = r->resid.empty() ? 0.0 : *std::max_element(r->resid.begin(), r->resid.end());
    info->batch_requests = r->batch_requests;
    info->batch_rhs = r->batch_rhs;
    info->cache_hit = r->cache_hit;
    info->slices = r->slices;
  }
  if (r->status != BMI_OK) bmi::set_last_error(r->error);
  return BMI_OK;
}

bmi_status bmi_server_fetch(bmi_server* srv, int64_t id, double* x) {
  if (!srv) return BMI_INVALID_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(id);
  if (it == srv->reqs.end() || it->second->state != 2) {
    bmi::set_last_error("request unknown or not finished");
    return BMI_INVALID_ARG;
  }
  if (x && !it->second->x.empty()) std::memcpy(x, it->second->x.data(), it->second->x.size() * sizeof(double));
  srv->reqs.erase(it);
  return BMI_OK;
}

bmi_status bmi_server_stats(bmi_server* srv, int64_t out[6]) {
  if (!srv || !out) return BMI_INVALID_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  out[0] = srv->submitted;
  out[1] = srv->finished;
  out[2] = static_cast<int64_t>(srv->queue.size());
  out[3] = srv->active;
  out[4] = static_cast<int64_t>(srv->ops.size());
  out[5] = srv->used_bytes;
  return BMI_OK;
}

bmi_status bmi_server_destroy(bmi_server* srv) {
  if (!srv) return BMI_OK;
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv.wait(lk, [&] { return srv->active == 0; });  // running batches finish first
    srv->stop = true;
    srv->cv.notify_all();
  }
  if (srv->scheduler.joinable()) srv->scheduler.join();
  if (srv->device.joinable()) srv->device.join();
  delete srv;
  return BMI_OK;
}

}  // extern "C"
