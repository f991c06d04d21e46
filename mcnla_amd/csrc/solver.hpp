// Host-side solver internals shared by the C ABI (capi.cpp) and the serving loop (serve.cpp).
#pragma once

#include <cstddef>
#include <cstdint>

#include "common.hpp"

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

}  // namespace mcnla
