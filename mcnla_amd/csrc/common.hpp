// Shared host/device helpers of libmcnla_amd (gfx950 / CDNA4 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mcnla_amd/isvd_capi.h"

namespace mcnla {

// ------------------------------------------------------------------------------------------------
// Error handling: every host entry point returns an mcnla_status; the message of the last
// failure is kept per thread.
void set_error(const std::string& msg);
const char* get_error();

struct Error {
  mcnla_status code;
  std::string msg;
};

#define MCNLA_HIP(expr)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      throw ::mcnla::Error{MCNLA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) \
                                              + " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"}; \
    }                                                                                       \
  } while (0)

#define MCNLA_REQUIRE(cond, code, text)                          \
  do {                                                           \
    if (!(cond)) throw ::mcnla::Error{(code), std::string(text)}; \
  } while (0)

// Launch-error check after every kernel launch.
#define MCNLA_LAUNCHED() MCNLA_HIP(hipGetLastError())

__host__ __device__ constexpr int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------------------------------------------
// Vector types used with the MFMA builtins.
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;

// Number of XCDs of an MI355X; workgroups are dispatched round-robin over them.
constexpr int kXcds = 8;

// Remaps a linear workgroup id so that consecutive *logical* ids land on the same XCD: the
// hardware sends workgroup b to XCD b % 8, so logical = (b % 8) * per_xcd + b / 8.  Ids beyond
// the last full round keep a valid (bijective) mapping.
__device__ __forceinline__ int64_t xcd_swizzle(int64_t b, int64_t nblocks) {
  const int64_t per = nblocks / kXcds;
  const int64_t full = per * kXcds;
  if (b >= full) return b;
  return (b % kXcds) * per + b / kXcds;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace mcnla
