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

// Device bytes of a worker's KN-group flags (2 kMaxKnGroup ints, one 256-byte line).
constexpr size_t kFlagBytes = 256;
static_assert(2 * kMaxKnGroup * sizeof(int) <= kFlagBytes, "flag words fit");

// Iterations a worker's KN group runs between admissions and completions.
constexpr int kChunk = 8;

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
  int bad = 0;  // the solve's hazard flag, read back on its stream
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
  int* flags = nullptr;  // per worker: 2 kMaxKnGroup ints of device memory for its KN group
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
  uint64_t batches = 0;  // admissions (a batch of one or more requests each)
  // KN groups: chunks run, members summed over chunks, wall seconds of the chunks
  uint64_t chunks = 0, chunk_members = 0;
  double chunk_seconds = 0.0;
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

  mcnla_solver prepare(Req* q, hipStream_t st, SolveRes* res, const void** A, int64_t* lda);
  void collect(Req* q, mcnla_solver s, hipStream_t st);
  void run(Req* q, hipStream_t st, SolveRes* res);
  void finish_locked(Req* q);
  void worker(int index);
  void worker_kn(hipStream_t st, SolveRes* res, int* flags, int cap);
};

// Makes q's A visible to a solve on `st` (after its ready event; a host A is copied behind the
// solver's workspaces) and creates its solver in its workspace region.
mcnla_solver mcnla_server_s::prepare(Req* q, hipStream_t st, SolveRes* res, const void** A,
                                     int64_t* lda) {
  char* ws = base + q->ws_off;
  const mcnla_params& p = q->r.params;
  *A = q->r.A;
  *lda = q->r.lda;
  if (q->r.ready_event)
    MCNLA_HIP(hipStreamWaitEvent(st, static_cast<hipEvent_t>(q->r.ready_event), 0));
  if (q->r.a_on_host) {
    // Dense copy of A after the solver's workspaces: lines of `len` elements, `count` of them.
    const size_t esz = q->r.a_dtype == MCNLA_F32 ? sizeof(float) : sizeof(double);
    const bool cm = q->r.a_order == MCNLA_COL_MAJOR;
    const int64_t len = cm ? p.nrow : p.ncol, count = cm ? p.ncol : p.nrow;
    char* dA = ws + q->solver_bytes;
    MCNLA_HIP(hipMemcpy2DAsync(dA, len * esz, q->r.A, q->r.lda * esz, len * esz, count,
                               hipMemcpyHostToDevice, st));
    *A = dA;
    *lda = len;
  }
  return solver_create_in(&p, &q->cfg, ws, q->solver_bytes, res);
}

// Enqueues the copies of a solved request's results into its result region.
void mcnla_server_s::collect(Req* q, mcnla_solver s, hipStream_t st) {
  char* rs = base + q->res_off;
  const mcnla_params& p = q->r.params;
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
}

// Solves one admitted request on `st` (mcnla_solver_solve) and copies its results.
void mcnla_server_s::run(Req* q, hipStream_t st, SolveRes* res) {
  mcnla_solver s = nullptr;
  try {
    const void* A = nullptr;
    int64_t lda = 0;
    s = prepare(q, st, res, &A, &lda);
    q->status = mcnla_solver_solve(s, A, lda, q->r.a_dtype, q->r.a_order, st);
    if (q->status != MCNLA_OK) q->error = mcnla_last_error();
    if (q->status == MCNLA_OK || q->status == MCNLA_ERR_NOT_CONVERGED) {
      collect(q, s, st);
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

// Bookkeeping of a finished request (its workspace region goes back to the arena).
void mcnla_server_s::finish_locked(Req* q) {
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

// Iteration-level batching of KN solves.  The worker keeps a group of up to `cap` solves whose KN
// iterations run together on its stream `st` (KnBatch: one launch per kernel per iteration for
// the whole group) and cycles:
//   - admit requests into free places of the group; each runs its sketch and orthogonalizer
//     stages on the worker's second stream `aux` (beside the group's iterations) and joins the
//     group at its iteration 0 once those have finished;
//   - run kChunk iterations of the group and read its flags (synchronizes `st`);
//   - hand the members that stopped to `aux` for the rest of their solve (coefficients, basis,
//     former, result copies); each completes when `aux` has passed its work.
// Nothing waits on the host while the group has members to iterate.  Every member computes
// exactly what it computes alone.
void mcnla_server_s::worker_kn(hipStream_t st, SolveRes* res, int* flags, int cap) {
  struct Live {
    Req* q = nullptr;
    mcnla_solver s = nullptr;
    const void* A = nullptr;
    int64_t lda = 0;
    SolveCtx cx;
    int slot = -1;  // place in the KN group
    int ev = -1;    // event: stages 1-2 done (joining) / results copied (finishing)
    bool failed = false;
  };
  hipStream_t aux = nullptr;
  hipEvent_t chunk_done = nullptr;
  std::vector<hipEvent_t> evs;  // one per member in flight: at most cap joining or live + cap
                                // finishing
  std::vector<int> free_ev;
  try {
    MCNLA_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    MCNLA_HIP(hipEventCreateWithFlags(&chunk_done, hipEventDisableTiming));
    for (int i = 0; i < 2 * cap; ++i) {
      hipEvent_t e = nullptr;
      MCNLA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      evs.push_back(e);
      free_ev.push_back(i);
    }
  } catch (const Error&) {
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    if (chunk_done) (void)hipEventDestroy(chunk_done);
    if (aux) (void)hipStreamDestroy(aux);
    aux = nullptr;
  }
  KnBatch kb(flags);
  uint64_t my_chunks = 0, my_members = 0;  // flushed into the server's totals under its lock
  double my_seconds = 0.0;
  std::vector<Live> joining, live, finishing;
  std::vector<Req*> fresh;
  auto done_event = [&](const Live& L) {
    const hipError_t e = hipEventQuery(evs[L.ev]);
    if (e == hipErrorNotReady) return false;
    (void)hipGetLastError();
    return true;
  };
  // completes a finishing member (its work on `aux` has passed, or it failed)
  auto complete = [&](Live& L) {
    if (!L.failed) {
      L.q->status = solve_status(L.s, L.cx, L.q->bad);
      if (L.q->status != MCNLA_OK) L.q->error = mcnla_last_error();
    }
    if (L.s) mcnla_solver_destroy(L.s);
    if (L.ev >= 0) free_ev.push_back(L.ev);
    std::lock_guard<std::mutex> lk(mu);
    finish_locked(L.q);
  };
  auto fail = [&](Live& L, mcnla_status code, const std::string& msg) {
    L.q->status = code;
    L.q->error = msg;
    L.failed = true;
    if (L.slot >= 0) kb.drop(L.slot);
    L.slot = -1;
  };
  for (;;) {
    // 1. admission
    fresh.clear();
    {
      std::unique_lock<std::mutex> lk(mu);
      chunks += my_chunks;
      chunk_members += my_members;
      chunk_seconds += my_seconds;
      my_chunks = my_members = 0;
      my_seconds = 0.0;
      const bool idle = joining.empty() && live.empty() && finishing.empty();
      if (idle) {
        cv_work.wait(lk, [&] {
          if (stop) return true;
          Req* q = (aux && !free_ev.empty()) ? pick_locked() : nullptr;
          if (q) fresh.push_back(q);
          return q != nullptr;
        });
        if (fresh.empty()) break;  // stopping, nothing in flight
      }
      while (!stop && aux && static_cast<int>(joining.size() + live.size() + fresh.size()) < cap &&
             free_ev.size() > fresh.size()) {
        Req* q = pick_locked();
        if (!q) break;
        fresh.push_back(q);
      }
      if (!fresh.empty()) {
        const Clock::time_point now = Clock::now();
        for (Req* q : fresh) {
          q->state = State::kRunning;
          q->t_admit = now;
        }
        running += static_cast<int>(fresh.size());
        max_running = std::max(max_running, running);
        ++batches;
      }
    }
    bool progressed = !fresh.empty();
    for (Req* q : fresh) {  // stages 1-2 on aux
      Live L;
      L.q = q;
      L.ev = free_ev.back();
      free_ev.pop_back();
      try {
        L.s = prepare(q, aux, res, &L.A, &L.lda);
        solve_check(L.s, L.A, L.lda, q->r.a_dtype, q->r.a_order);
        solve_begin(L.s, L.A, L.lda, q->r.a_dtype, q->r.a_order, aux, L.cx, false);
      } catch (const Error& e) {
        fail(L, e.code, e.msg);
      } catch (const std::exception& e) {
        fail(L, MCNLA_ERR_INVALID, e.what());
      }
      (void)hipEventRecord(evs[L.ev], aux);
      joining.push_back(L);
    }
    // 2. members whose stages 1-2 are done join the group (wait for the oldest when nothing else
    // can run)
    if (live.empty() && !joining.empty()) (void)hipEventSynchronize(evs[joining.front().ev]);
    for (size_t i = 0; i < joining.size();) {
      Live& L = joining[i];
      if (!done_event(L)) {
        ++i;
        continue;
      }
      if (!L.failed) {
        try {
          MCNLA_HIP(hipStreamWaitEvent(st, evs[L.ev], 0));
          L.slot = kb.add(solver_kn_integrator(L.s), st);
        } catch (const Error& e) {
          fail(L, e.code, e.msg);
        }
      }
      if (L.failed) {
        (void)hipStreamSynchronize(aux);
        if (res->side) (void)hipStreamSynchronize(res->side);
        complete(L);
      } else {
        live.push_back(L);
      }
      joining[i] = joining.back();
      joining.pop_back();
      progressed = true;
    }
    // 3. a chunk of the group's iterations and its flags (synchronizes st)
    if (!live.empty()) {
      progressed = true;
      try {
        const Clock::time_point c0 = Clock::now();
        kb.advance(kChunk, st);
        ++my_chunks;
        my_members += live.size();
        my_seconds += seconds(c0, Clock::now());
        MCNLA_HIP(hipEventRecord(chunk_done, st));
        MCNLA_HIP(hipStreamWaitEvent(aux, chunk_done, 0));
      } catch (const Error& e) {
        for (Live& L : live) fail(L, e.code, e.msg);
      }
    }
    // 4. members that stopped: the rest of their solve on aux
    for (size_t i = 0; i < live.size();) {
      Live& L = live[i];
      if (!L.failed && kb.running(L.slot)) {
        ++i;
        continue;
      }
      if (!L.failed) {
        try {
          const int dt = L.q->r.a_dtype, od = L.q->r.a_order;
          L.cx.conv = kb.converged(L.slot);
          L.cx.ok = kb.numeric_ok(L.slot);
          L.q->iterations = kb.iterations(L.slot);
          solver_set_integration(L.s, L.q->iterations, L.cx.conv);
          kb.finish(L.slot, solver_coef(L.s), aux);
          L.slot = -1;
          solve_finish(L.s, L.A, L.lda, dt, od, aux, L.cx, false);
          read_bad(L.s, &L.q->bad, aux);
          collect(L.q, L.s, aux);  // a failed solve's copies are never handed out
        } catch (const Error& e) {
          fail(L, e.code, e.msg);
        }
      }
      (void)hipEventRecord(evs[L.ev], aux);
      finishing.push_back(L);
      live[i] = live.back();
      live.pop_back();
    }
    // 5. complete the finishing members that aux has passed (wait for the oldest when nothing
    // else can run)
    if (live.empty() && joining.empty() && !finishing.empty())
      (void)hipEventSynchronize(evs[finishing.front().ev]);
    for (size_t i = 0; i < finishing.size();) {
      Live& L = finishing[i];
      if (!done_event(L)) {
        ++i;
        continue;
      }
      if (L.failed) {
        (void)hipStreamSynchronize(aux);
        if (res->side) (void)hipStreamSynchronize(res->side);
      }
      complete(L);
      finishing[i] = finishing.back();
      finishing.pop_back();
      progressed = true;
    }
    if (progressed) {
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
  if (aux) {
    (void)hipStreamSynchronize(aux);
    (void)hipStreamDestroy(aux);
  }
  for (hipEvent_t e : evs) (void)hipEventDestroy(e);
  if (chunk_done) (void)hipEventDestroy(chunk_done);
}

// A worker owns its stream and the HIP objects its solves record into (SolveRes) for its whole
// life, so that admitting, running and finishing a request creates and destroys no HIP stream or
// event.
void mcnla_server_s::worker(int index) {
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
  // KN solves with batch_max > 1 run in the worker's KN group; everything else one at a time.
  const int kind = cfg.solver.integrator;
  const bool kn = kind == MCNLA_INT_KN || kind == MCNLA_INT_ROW_BLOCK_KN ||
                  kind == MCNLA_INT_ROW_BLOCK_GRAMIAN_KN;
  if (st && kn && cfg.batch_max > 1) {
    worker_kn(st, &res, flags + index * (kFlagBytes / sizeof(int)), cfg.batch_max);
  } else {
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
        ++batches;
      }
      run(q, st, &res);
      {
        std::lock_guard<std::mutex> lk(mu);
        finish_locked(q);
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
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
  cfg->batch_max = 1;
}

mcnla_status mcnla_server_create(const mcnla_server_config* cfg, mcnla_server* out) {
  return serve_guard([&] {
    MCNLA_REQUIRE(cfg && out, MCNLA_ERR_INVALID, "null argument");
    MCNLA_REQUIRE(cfg->slots >= 1 && cfg->slots <= 16, MCNLA_ERR_INVALID, "slots must be 1..16");
    MCNLA_REQUIRE(cfg->hbm_budget >= kCarveAlign, MCNLA_ERR_INVALID, "hbm_budget too small");
    MCNLA_REQUIRE(cfg->max_bypass >= 0, MCNLA_ERR_INVALID, "max_bypass must be >= 0");
    MCNLA_REQUIRE(cfg->batch_max >= 1 && cfg->batch_max <= MCNLA_MAX_BATCH, MCNLA_ERR_INVALID,
                  "batch_max must be 1..16");
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
    size_t foff = 0;  // the workers' flag words, at the front of the arena
    MCNLA_REQUIRE(s->arena.take(static_cast<size_t>(s->cfg.slots) * kFlagBytes, &foff),
                  MCNLA_ERR_NOMEM, "hbm_budget too small");
    s->flags = reinterpret_cast<int*>(s->base + foff);
    for (int i = 0; i < s->cfg.slots; ++i)
      s->workers.emplace_back([srv = s.get(), i] { srv->worker(i); });
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
    out->batches = s->batches;
    out->kn_chunks = s->chunks;
    out->kn_group_mean = s->chunks ? static_cast<double>(s->chunk_members) / s->chunks : 0.0;
    out->kn_chunk_mean_s = s->chunks ? s->chunk_seconds / s->chunks : 0.0;
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
    s->batches = 0;
    s->chunks = s->chunk_members = 0;
    s->chunk_seconds = 0.0;
    s->max_running = s->running;
    s->latency.clear();
    s->any_submit = false;
    s->arena.reset_peak();
    return MCNLA_OK;
  });
}

}  // extern "C"
