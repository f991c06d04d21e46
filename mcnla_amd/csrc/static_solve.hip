// mcnla_solve_static: a fused solve whose every device workspace is carved from a file-scope
// device array, so the call allocates no device memory.
//
// The array (`scratch`) exists once per GPU the library runs on and lives until the process ends.
// In host code its name is only a host-side stand-in: the device address comes from
// hipGetSymbolAddress, looked up once per device (with that device current) and kept per device;
// only that address is ever handed to the solver and its kernels.  Every call on one GPU carves
// the same array, so calls that overlap on one GPU race (call it from one host thread per GPU, or
// serialize the calls); sequential calls are independent because a solve initializes every flag,
// counter and accumulator it reads (split-K partials are written before they are summed, the
// integrator's state block is reset by its first kernel, the status flag is cleared per solve).
#include <memory>
#include <mutex>
#include <string>

#include "common.hpp"
#include "solver.hpp"

namespace mcnla {
namespace {

constexpr size_t kStaticBytes = MCNLA_STATIC_WORKSPACE_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kStaticBytes];

struct PerDevice {
  void* ptr = nullptr;  // device address of `scratch` on this GPU
  SolveRes res;         // side stream and events the solves on this GPU record into
};

std::mutex g_mu;
PerDevice g_dev[kMaxDevices];

// The scratch address and solve resources of the current device, set up on first use.
PerDevice& current_device() {
  int dev = 0;
  MCNLA_HIP(hipGetDevice(&dev));
  MCNLA_REQUIRE(dev >= 0 && dev < kMaxDevices, MCNLA_ERR_NO_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  PerDevice& d = g_dev[dev];
  if (!d.ptr) {
    void* p = nullptr;
    MCNLA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MCNLA_REQUIRE(p != nullptr && reinterpret_cast<uintptr_t>(p) % kCarveAlign == 0,
                  MCNLA_ERR_HIP, "static workspace has no aligned device address");
    d.res.create();
    d.ptr = p;
  }
  return d;
}

template <class F>
mcnla_status static_guard(F&& f) {
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

struct SolverDeleter {
  void operator()(mcnla_solver s) const { mcnla_solver_destroy(s); }
};

}  // namespace
}  // namespace mcnla

using namespace mcnla;

extern "C" mcnla_status mcnla_solve_static(const mcnla_params* params,
                                           const mcnla_solver_config* cfg, const void* A,
                                           int64_t lda, int a_dtype, int a_order, double* S,
                                           double* U, double* V, double* Qbar,
                                           int64_t* iterations, void* stream) {
  return static_guard([&] {
    MCNLA_REQUIRE(params && cfg, MCNLA_ERR_INVALID, "null argument");
    MCNLA_REQUIRE(cfg->graph == 0, MCNLA_ERR_INVALID,
                  "graph recording allocates; the static-workspace solve runs eagerly");
    const size_t need = solver_required_bytes(params, cfg);  // validates params and cfg
    MCNLA_REQUIRE(need <= kStaticBytes, MCNLA_ERR_NOMEM,
                  "the solve needs " + std::to_string(need) +
                      " bytes of workspace; the static workspace holds " +
                      std::to_string(kStaticBytes));
    PerDevice& d = current_device();
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::unique_ptr<mcnla_solver_s, SolverDeleter> s(
        solver_create_in(params, cfg, d.ptr, kStaticBytes, &d.res));
    const mcnla_status status = mcnla_solver_solve(s.get(), A, lda, a_dtype, a_order, stream);
    if (status != MCNLA_OK && status != MCNLA_ERR_NOT_CONVERGED) return status;
    const std::string warning = status == MCNLA_OK ? "" : mcnla_last_error();
    const double *Sd = nullptr, *Ud = nullptr, *Vd = nullptr, *Qd = nullptr;
    if (mcnla_solver_results(s.get(), &Sd, &Ud, &Vd, &Qd) != MCNLA_OK)
      throw Error{MCNLA_ERR_STATE, mcnla_last_error()};
    const size_t m = static_cast<size_t>(params->nrow), n = static_cast<size_t>(params->ncol);
    const size_t k = static_cast<size_t>(params->rank);
    const size_t l = static_cast<size_t>(params->rank + params->over_rank);
    auto out = [&](double* dst, const double* src, size_t count) {
      if (dst && src) MCNLA_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                                               hipMemcpyDeviceToDevice, st));
    };
    out(S, Sd, k);
    out(U, Ud, m * k);
    out(V, Vd, n * k);
    out(Qbar, Qd, m * l);
    int64_t it = 0;
    mcnla_solver_iteration(s.get(), &it);
    if (iterations) *iterations = it;
    // the copies read the array that the next call on this GPU overwrites
    MCNLA_HIP(hipStreamSynchronize(st));
    if (status != MCNLA_OK) set_error(warning);
    return status;
  });
}
