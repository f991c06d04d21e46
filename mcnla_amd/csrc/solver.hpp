// Host-side solver internals shared by the C ABI (capi.cpp) and the serving loop (serve.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace mcnla {

// Alignment of every buffer carved from a region (a multiple of the 16-byte DMA granule and of
// the 256-byte HBM channel interleave).
constexpr size_t kCarveAlign = 256;

constexpr size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// A region that device buffers are carved from, front to back.  A counting region has no memory:
// carving only adds up the aligned sizes, which sizes a solver without allocating.
struct Carve {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  bool counting = false;
};

// Makes `c` the region that device buffers created on this thread draw from, until destroyed.
class CarveScope {
 public:
  explicit CarveScope(Carve* c);
  ~CarveScope();
  CarveScope(const CarveScope&) = delete;
  CarveScope& operator=(const CarveScope&) = delete;

 private:
  Carve* prev_;
};

Carve* current_carve();

// The HIP objects a fused solve records into: the side stream of the first block's
// eigen-solve, its fork/join events, the stage marks and the begin/end events of the two timed
// GEMM launches.  A stand-alone solver creates its own set on its first solve.  A long-lived
// owner (a server worker, the static-workspace entry point) creates one set and lends it to
// every solver it runs, so that creating, running and destroying a solver creates and destroys
// no HIP stream or event.  Solves that share one set must not overlap.
struct SolveRes {
  static constexpr int kMarks = 5;  // start + one per stage
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t marks[kMarks] = {};
  hipEvent_t sketch_beg = nullptr, sketch_end = nullptr, gram_beg = nullptr, gram_end = nullptr;
  hipEvent_t former_beg = nullptr, former_end = nullptr;
  bool ready() const { return side != nullptr; }
  void create();   // on the current device; throws Error (and holds nothing) on failure
  void destroy();  // idempotent, never throws
};

// Device bytes a solver of this shape and configuration holds (every workspace, Omega and the
// results; not A).  Allocates nothing and needs no device.  Throws Error.
size_t solver_required_bytes(const mcnla_params* params, const mcnla_solver_config* cfg);

// Creates a solver whose device buffers all lie in [base, base + bytes); `bytes` must be at least
// solver_required_bytes().  The solver never allocates afterwards.  With `res` (created, and
// outliving the solver) the solver records into those HIP objects instead of its own.
// Throws Error.
mcnla_solver solver_create_in(const mcnla_params* params, const mcnla_solver_config* cfg,
                              void* base, size_t bytes, SolveRes* res = nullptr);

// The phases of a fused solve, for callers that run the integration themselves (the batch solve
// and the serving loop's iteration-level batching).  The solve of mcnla_solver_solve is
// solve_begin, the integration of s->Bs into solver_coef(s), solver_set_integration,
// solve_finish, read_bad, a synchronization and solve_status.
struct SolveCtx {
  const double* Ar = nullptr;  // Gram-of-A route: A as row-major FP64
  int64_t ldar = 0;
  bool rotate = false;         // Cholesky transforms + the first block's eigen rotation
  bool conv = false, ok = true;
};
// Checks A against the solver's shape (throws Error).
void solve_check(mcnla_solver s, const void* A, int64_t lda, int a_dtype, int a_order);
// Stages 1-2 on `st`: the sketch Y = A Omega (or H = A'A), G = Y'Y, the per-block transforms T
// and Bs = T'GT.  `timed`: record the stage marks.
void solve_begin(mcnla_solver s, const void* A, int64_t lda, int a_dtype, int a_order,
                 hipStream_t st, SolveCtx& cx, bool timed);
// Stages 3-4 after the integration left the coefficients in solver_coef(s): Qbar = Y (T coef)
// (or A (Omega T coef)), then the former.
void solve_finish(mcnla_solver s, const void* A, int64_t lda, int a_dtype, int a_order,
                  hipStream_t st, const SolveCtx& cx, bool timed);
// Enqueues the read-back of the solve's hazard flag into *bad (the caller synchronizes).
void read_bad(mcnla_solver s, int* bad, hipStream_t st);
// The solve's status once its stream has finished (sets the thread's error message).
mcnla_status solve_status(const mcnla_solver_s* s, const SolveCtx& cx, int bad);
bool solver_is_kn(mcnla_solver s);
// The solver's KN integration on its Bs and workspace.
GramIntegrator solver_kn_integrator(mcnla_solver s);
double* solver_coef(mcnla_solver s);
void solver_set_integration(mcnla_solver s, int64_t iterations, bool converged);

// The batch solve of mcnla_solver_solve_batch on stream `st` (synchronized on return); with
// `errors`, member i's error text (empty when status[i] is OK).  Throws Error for invalid
// arguments (before anything is enqueued) or a HIP failure.
void solve_batch(int count, const mcnla_solver* solvers, const void* const* A, const int64_t* lda,
                 const int* a_dtype, const int* a_order, mcnla_status* status,
                 std::vector<std::string>* errors, hipStream_t st);

}  // namespace mcnla
