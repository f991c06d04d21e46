// Serving loop of libmcnla_amd: continuous batching of iSVD requests on one GPU.
//
// One device arena, allocated when the server is created, holds everything a request needs on
// the device: its solver's workspaces (Y, the Gram matrix, Omega, integrator state, former
// buffers), a copy of A when A comes from the host, and its results.  Worker threads, each with
// its own HIP stream, take requests from a FIFO queue as soon as a worker is free and the arena has
// a free region of the request's size; a solve is created inside that region, so the serving
// loop never allocates device memory.  The workspace region is returned when the solve finishes;
// the results stay until the caller releases them.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "solver.hpp"

namespace mcnla {
namespace {

using Clock = std::chrono::steady_clock;

double seconds(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

// First-fit allocator over the offsets of the device arena (host bookkeeping only).  Free ranges
// are kept by offset and merged with their neighbours when returned.
class Arena {
 public:
  explicit Arena(size_t cap) : cap_(cap) {
    if (cap_) free_[0] = cap_;
  }
  static size_t round(size_t n) { return align_up(std::max<size_t>(n, 1), kCarveAlign); }
  bool take(size_t n, size_t* off) {
    n = round(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n, at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += n;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void give(size_t off, size_t n) {
    n = round(n);
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t cap() const { return cap_; }
  size_t peak() const { return peak_; }
  void reset_peak() { peak_ = used_; }

 private:
  size_t cap_, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

enum class State { kPending, kRunning, kDone };

// Offsets of the results inside a request's result region.
struct ResultLayout {
  size_t s = 0, u = 0, v = 0, q = 0, bytes = 0;
  bool has_v = false;
};

ResultLayout result_layout(const mcnla_params& p, const mcnla_solver_config& c) {
  ResultLayout r;
  const size_t k = static_cast<size_t>(p.rank), l = static_cast<size_t>(p.rank + p.over_rank);
  r.has_v = c.jobv != 0 && c.former != MCNLA_FORM_ROW_BLOCK_SYMMETRIC;
  size_t at = 0;
  auto put = [&](size_t bytes) {
    const size_t off = at;
    at = align_up(at + bytes, kCarveAlign);
    return off;
  };
  r.s = put(k * sizeof(double));
  r.u = put(static_cast<size_t>(p.nrow) * k * sizeof(double));
  if (r.has_v) r.v = put(static_cast<size_t>(p.ncol) * k * sizeof(double));
  r.q = put(static_cast<size_t>(p.nrow) * l * sizeof(double));
  r.bytes = at;
  return r;
}

struct Req {
  uint64_t id = 0;
  mcnla_request r{};
  mcnla_solver_config cfg{};
  size_t solver_bytes = 0, a_bytes = 0, ws_bytes = 0;
  ResultLayout res;
  size_t ws_off = 0, res_off = 0;
  State state = State::kPending;
  mcnla_status status = MCNLA_OK;
  std::string error;
  int64_t iterations = 0;
  Clock::time_point t_submit, t_admit, t_done;
};

size_t host_a_bytes(const mcnla_request& r) {
  if (!r.a_on_host) return 0;
  const size_t esz = r.a_dtype == MCNLA_F32 ? sizeof(float) : sizeof(double);
  return align_up(static_cast<size_t>(r.params.nrow) * static_cast<size_t>(r.params.ncol) * esz,
                  kCarveAlign);
}

}  // namespace
}  // namespace mcnla

using namespace mcnla;

struct mcnla_server_s {
  mcnla_server_config cfg{};
  int device = 0;
  char* base = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  Arena arena{0};
  std::deque<Req*> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  std::deque<uint64_t> done_q;
  std::vector<std::thread> workers;
  bool stop = false;
  uint64_t next_id = 1;
  int head_bypassed = 0;
  int running = 0;
  // statistics since the last reset
  uint64_t submitted = 0, completed = 0, failed = 0;
  int max_running = 0;
  std::vector<double> latency;
  bool any_submit = false;
  Clock::time_point t_first, t_last;

  // Reserves both regions of `q`; false (and nothing held) if either does not fit.
  bool reserve(Req* q) {
    if (!arena.take(q->ws_bytes, &q->ws_off)) return false;
    if (!arena.take(q->res.bytes, &q->res_off)) {
      arena.give(q->ws_off, q->ws_bytes);
      return false;
    }
    return true;
  }
  // The next admissible request: the queue head, or one of the `max_bypass` requests behind it
  // while the head has been passed over fewer than `max_bypass` times.
  Req* pick_locked() {
    int scanned = 0;
    for (auto it = pending.begin(); it != pending.end(); ++it, ++scanned) {
      const bool head = (it == pending.begin());
      if (!head && (scanned > cfg.max_bypass || head_bypassed >= cfg.max_bypass)) break;
      Req* q = *it;
      if (!reserve(q)) continue;
      head_bypassed = head ? 0 : head_bypassed + 1;
      pending.erase(it);
      return q;
    }
    return nullptr;
  }
  // client threads blocked in mcnla_server_wait / next_done; destroy waits until none is left
  int waiters = 0;
  std::condition_variable cv_left;

  void run(Req* q, hipStream_t st, SolveRes* res);
  void worker();
};

void mcnla_server_s::run(Req* q, hipStream_t st, SolveRes* res) {
  char* ws = base + q->ws_off;
  char* rs = base + q->res_off;
  mcnla_solver s = nullptr;
  try {
    const mcnla_params& p = q->r.params;
    const void* A = q->r.A;
    int64_t lda = q->r.lda;
    if (q->r.ready_event)
      MCNLA_HIP(hipStreamWaitEvent(st, static_cast<hipEvent_t>(q->r.ready_event), 0));
    if (q->r.a_on_host) {
      // Dense copy of A after the solver's workspaces: lines of `len` elements, `count` of them.
      const size_t esz = q->r.a_dtype == MCNLA_F32 ? sizeof(float) : sizeof(double);
      const bool cm = q->r.a_order == MCNLA_COL_MAJOR;
      const int64_t len = cm ? p.nrow : p.ncol, count = cm ? p.ncol : p.nrow;
      char* dA = ws + q->solver_bytes;
      MCNLA_HIP(hipMemcpy2DAsync(dA, len * esz, A, lda * esz, len * esz, count,
                                 hipMemcpyHostToDevice, st));
      A = dA;
      lda = len;
    }
    s = solver_create_in(&p, &q->cfg, ws, q->solver_bytes, res);
    q->status = mcnla_solver_solve(s, A, lda, q->r.a_dtype, q->r.a_order, st);
    if (q->status != MCNLA_OK) q->error = mcnla_last_error();
    if (q->status == MCNLA_OK || q->status == MCNLA_ERR_NOT_CONVERGED) {
      const double *S = nullptr, *U = nullptr, *V = nullptr, *Qb = nullptr;
      if (mcnla_solver_results(s, &S, &U, &V, &Qb) != MCNLA_OK)
        throw Error{MCNLA_ERR_STATE, mcnla_last_error()};
      const size_t k = static_cast<size_t>(p.rank), l = static_cast<size_t>(p.rank + p.over_rank);
      MCNLA_HIP(hipMemcpyAsync(rs + q->res.s, S, k * sizeof(double), hipMemcpyDeviceToDevice, st));
      MCNLA_HIP(hipMemcpyAsync(rs + q->res.u, U, p.nrow * k * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
      if (q->res.has_v && V)
        MCNLA_HIP(hipMemcpyAsync(rs + q->res.v, V, p.ncol * k * sizeof(double),
                                 hipMemcpyDeviceToDevice, st));
      MCNLA_HIP(hipMemcpyAsync(rs + q->res.q, Qb, p.nrow * l * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
      MCNLA_HIP(hipStreamSynchronize(st));
      mcnla_solver_iteration(s, &q->iterations);
    }
  } catch (const Error& e) {
    q->status = e.code;
    q->error = e.msg;
  } catch (const std::exception& e) {
    q->status = MCNLA_ERR_INVALID;
    q->error = e.what();
  }
  // Whatever happened, both streams are idle before the solver goes and the workspace region is
  // handed back (a failed solve may have left work queued behind its error).
  (void)hipStreamSynchronize(st);
  if (res->side) (void)hipStreamSynchronize(res->side);
  if (s) mcnla_solver_destroy(s);
}

// A worker owns its stream and the HIP objects its solves record into (SolveRes) for its whole
// life, so that admitting, running and finishing a request creates and destroys no HIP stream or
// event.
void mcnla_server_s::worker() {
  (void)hipSetDevice(device);
  hipStream_t st = nullptr;
  SolveRes res;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    st = nullptr;
  } else {
    try {
      res.create();
    } catch (const Error&) {
      (void)hipStreamDestroy(st);
      st = nullptr;
    }
  }
  for (;;) {
    Req* q = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || (st && (q = pick_locked()) != nullptr); });
      if (!q) break;
      q->state = State::kRunning;
      q->t_admit = Clock::now();
      ++running;
      max_running = std::max(max_running, running);
    }
    run(q, st, &res);
    {
      std::lock_guard<std::mutex> lk(mu);
      q->t_done = Clock::now();
      q->state = State::kDone;
      arena.give(q->ws_off, q->ws_bytes);
      --running;
      const bool ok = q->status == MCNLA_OK || q->status == MCNLA_ERR_NOT_CONVERGED;
      if (ok) {
        ++completed;
        latency.push_back(seconds(q->t_submit, q->t_done));
      } else {
        ++failed;
      }
      t_last = q->t_done;
      done_q.push_back(q->id);
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
  res.destroy();
  if (st) (void)hipStreamDestroy(st);
}

namespace {

template <class F>
mcnla_status serve_guard(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return MCNLA_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return MCNLA_ERR_INVALID;
  }
}

void fill_request(mcnla_server s, const mcnla_request* r, Req* q) {
  MCNLA_REQUIRE(r != nullptr, MCNLA_ERR_INVALID, "request is null");
  q->r = *r;
  q->cfg = s->cfg.solver;
  q->cfg.seed = r->seed;
  q->solver_bytes = solver_required_bytes(&r->params, &q->cfg);
  MCNLA_REQUIRE(r->A != nullptr, MCNLA_ERR_INVALID, "A is null");
  MCNLA_REQUIRE(r->a_dtype == MCNLA_F64 || r->a_dtype == MCNLA_F32, MCNLA_ERR_INVALID, "A dtype");
  MCNLA_REQUIRE(r->a_order == MCNLA_COL_MAJOR || r->a_order == MCNLA_ROW_MAJOR, MCNLA_ERR_INVALID,
                "A order");
  MCNLA_REQUIRE(r->lda >= (r->a_order == MCNLA_COL_MAJOR ? r->params.nrow : r->params.ncol),
                MCNLA_ERR_INVALID, "lda smaller than the leading dimension of A");
  q->a_bytes = host_a_bytes(*r);
  q->ws_bytes = q->solver_bytes + q->a_bytes;
  q->res = result_layout(r->params, q->cfg);
}

}  // namespace

extern "C" {

void mcnla_server_config_default(mcnla_server_config* cfg) {
  if (!cfg) return;
  mcnla_solver_config_default(&cfg->solver);
  cfg->slots = 8;
  cfg->hbm_budget = static_cast<size_t>(16) << 30;
  cfg->max_bypass = 8;
}

mcnla_status mcnla_server_create(const mcnla_server_config* cfg, mcnla_server* out) {
  return serve_guard([&] {
    MCNLA_REQUIRE(cfg && out, MCNLA_ERR_INVALID, "null argument");
    MCNLA_REQUIRE(cfg->slots >= 1 && cfg->slots <= 16, MCNLA_ERR_INVALID, "slots must be 1..16");
    MCNLA_REQUIRE(cfg->hbm_budget >= kCarveAlign, MCNLA_ERR_INVALID, "hbm_budget too small");
    MCNLA_REQUIRE(cfg->max_bypass >= 0, MCNLA_ERR_INVALID, "max_bypass must be >= 0");
    auto s = std::make_unique<mcnla_server_s>();
    s->cfg = *cfg;
    s->cfg.hbm_budget = align_up(cfg->hbm_budget, kCarveAlign);
    MCNLA_HIP(hipGetDevice(&s->device));
    void* p = nullptr;
    if (hipMalloc(&p, s->cfg.hbm_budget) != hipSuccess) {
      (void)hipGetLastError();
      throw Error{MCNLA_ERR_NOMEM, "cannot allocate the server arena of " +
                                       std::to_string(s->cfg.hbm_budget) + " bytes"};
    }
    s->base = static_cast<char*>(p);
    s->arena = Arena(s->cfg.hbm_budget);
    for (int i = 0; i < s->cfg.slots; ++i) s->workers.emplace_back([srv = s.get()] { srv->worker(); });
    *out = s.release();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_destroy(mcnla_server s) {
  return serve_guard([&] {
    if (!s) return MCNLA_OK;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    {
      // client threads still waiting on a request fail with MCNLA_ERR_STATE; none may touch the
      // server after this block
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_done.notify_all();
      s->cv_left.wait(lk, [&] { return s->waiters == 0; });
    }
    (void)hipFree(s->base);
    delete s;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_request_bytes(mcnla_server s, const mcnla_request* r, size_t* bytes) {
  return serve_guard([&] {
    MCNLA_REQUIRE(s && bytes, MCNLA_ERR_INVALID, "null argument");
    Req q;
    fill_request(s, r, &q);
    *bytes = Arena::round(q.ws_bytes) + Arena::round(q.res.bytes);
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_submit(mcnla_server s, const mcnla_request* r, uint64_t* id) {
  return serve_guard([&] {
    MCNLA_REQUIRE(s && id, MCNLA_ERR_INVALID, "null argument");
    auto q = std::make_unique<Req>();
    fill_request(s, r, q.get());
    MCNLA_REQUIRE(Arena::round(q->ws_bytes) + Arena::round(q->res.bytes) <= s->arena.cap(),
                  MCNLA_ERR_NOMEM, "request needs more device memory than the server budget");
    {
      std::lock_guard<std::mutex> lk(s->mu);
      MCNLA_REQUIRE(!s->stop, MCNLA_ERR_STATE, "server is shutting down");
      q->id = s->next_id++;
      q->t_submit = Clock::now();
      if (!s->any_submit) {
        s->any_submit = true;
        s->t_first = q->t_submit;
      }
      ++s->submitted;
      *id = q->id;
      s->pending.push_back(q.get());
      s->reqs.emplace(q->id, std::move(q));
    }
    s->cv_work.notify_all();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_wait(mcnla_server s, uint64_t id, double timeout_s, int* done) {
  return serve_guard([&] {
    MCNLA_REQUIRE(s && done, MCNLA_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    MCNLA_REQUIRE(s->reqs.count(id), MCNLA_ERR_INVALID, "unknown request id");
    // The request is looked up again on every wake-up: another thread may release it meanwhile.
    auto gone = [&] { return !s->reqs.count(id); };
    auto pred = [&] { return s->stop || gone() || s->reqs.at(id)->state == State::kDone; };
    ++s->waiters;
    if (timeout_s < 0)
      s->cv_done.wait(lk, pred);
    else
      s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    if (--s->waiters == 0) s->cv_left.notify_all();
    MCNLA_REQUIRE(!gone(), MCNLA_ERR_INVALID, "request was released while waiting");
    *done = s->reqs.at(id)->state == State::kDone ? 1 : 0;
    MCNLA_REQUIRE(*done || !s->stop, MCNLA_ERR_STATE, "server is shutting down");
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_next_done(mcnla_server s, double timeout_s, uint64_t* id) {
  return serve_guard([&] {
    MCNLA_REQUIRE(s && id, MCNLA_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] {
      while (!s->done_q.empty() && !s->reqs.count(s->done_q.front())) s->done_q.pop_front();
      return !s->done_q.empty();
    };
    auto pred = [&] { return ready() || s->stop; };
    ++s->waiters;
    if (timeout_s < 0)
      s->cv_done.wait(lk, pred);
    else
      s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    if (--s->waiters == 0) s->cv_left.notify_all();
    *id = 0;
    if (ready()) {
      *id = s->done_q.front();
      s->done_q.pop_front();
      return MCNLA_OK;
    }
    MCNLA_REQUIRE(!s->stop, MCNLA_ERR_STATE, "server is shutting down");
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_result(mcnla_server s, uint64_t id, mcnla_served* out) {
  return serve_guard([&] {
    MCNLA_REQUIRE(s && out, MCNLA_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    MCNLA_REQUIRE(it != s->reqs.end(), MCNLA_ERR_INVALID, "unknown request id");
    const Req* q = it->second.get();
    MCNLA_REQUIRE(q->state == State::kDone, MCNLA_ERR_STATE, "request has not finished");
    const bool ok = q->status == MCNLA_OK || q->status == MCNLA_ERR_NOT_CONVERGED;
    const char* rs = s->base + q->res_off;
    out->status = q->status;
    out->iterations = q->iterations;
    out->S = ok ? reinterpret_cast<const double*>(rs + q->res.s) : nullptr;
    out->U = ok ? reinterpret_cast<const double*>(rs + q->res.u) : nullptr;
    out->V = (ok && q->res.has_v) ? reinterpret_cast<const double*>(rs + q->res.v) : nullptr;
    out->Qbar = ok ? reinterpret_cast<const double*>(rs + q->res.q) : nullptr;
    out->queue_s = seconds(q->t_submit, q->t_admit);
    out->run_s = seconds(q->t_admit, q->t_done);
    out->latency_s = seconds(q->t_submit, q->t_done);
    out->error = q->error.c_str();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_release(mcnla_server s, uint64_t id) {
  return serve_guard([&] {
    MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "server is null");
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->reqs.find(id);
      MCNLA_REQUIRE(it != s->reqs.end(), MCNLA_ERR_INVALID, "unknown request id");
      MCNLA_REQUIRE(it->second->state == State::kDone, MCNLA_ERR_STATE, "request has not finished");
      s->arena.give(it->second->res_off, it->second->res.bytes);
      s->reqs.erase(it);
    }
    s->cv_work.notify_all();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_get_stats(mcnla_server s, mcnla_server_stats* out) {
  return serve_guard([&] {
    MCNLA_REQUIRE(s && out, MCNLA_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->max_running = s->max_running;
    out->peak_bytes = s->arena.peak();
    out->budget_bytes = s->arena.cap();
    out->p50_s = out->p99_s = out->mean_s = out->max_s = out->span_s = 0.0;
    if (!s->latency.empty()) {
      std::vector<double> v = s->latency;
      std::sort(v.begin(), v.end());
      // nearest-rank percentiles
      auto rank = [&](double q) {
        size_t r = static_cast<size_t>(std::ceil(q * static_cast<double>(v.size())));
        return v[std::min(v.size(), std::max<size_t>(r, 1)) - 1];
      };
      out->p50_s = rank(0.50);
      out->p99_s = rank(0.99);
      double sum = 0.0;
      for (double x : v) sum += x;
      out->mean_s = sum / static_cast<double>(v.size());
      out->max_s = v.back();
      out->span_s = seconds(s->t_first, s->t_last);
    }
    return MCNLA_OK;
  });
}

mcnla_status mcnla_server_reset_stats(mcnla_server s) {
  return serve_guard([&] {
    MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "server is null");
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted = s->completed = s->failed = 0;
    s->max_running = s->running;
    s->latency.clear();
    s->any_submit = false;
    s->arena.reset_peak();
    return MCNLA_OK;
  });
}

}  // extern "C"
