// Counter-based Gaussian random numbers (Philox4x32-10, Salmon et al., SC'11, + Box-Muller) and
// the synthetic low-rank test-matrix generator.
//
// Element e of a stream depends only on (seed, e): pair p = e / 2 uses counter (p, 0, 0, 0) and
// key (seed_lo, seed_hi); its four 32-bit outputs make two 53-bit uniforms u1 in (0,1],
// u2 in [0,1), and z = sqrt(-2 ln u1) * (cos, sin)(2 pi u2) gives elements 2p and 2p+1.  The
// result is independent of the launch shape, so every precision and every shape of a solve sees
// the same Omega.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace mcnla {

namespace {

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;
  constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(kM0, ctr.x), lo0 = kM0 * ctr.x;
    const uint32_t hi1 = __umulhi(kM1, ctr.z), lo1 = kM1 * ctr.z;
    ctr = U4{hi1 ^ ctr.y ^ k0, lo1, hi0 ^ ctr.w ^ k1, lo0};
    k0 += kW0;
    k1 += kW1;
  }
  return ctr;
}

__device__ __forceinline__ void gaussian_pair(uint64_t seed, uint64_t pair, double& z0,
                                              double& z1) {
  const U4 r = philox4x32_10(U4{static_cast<uint32_t>(pair), static_cast<uint32_t>(pair >> 32), 0u, 0u},
                             static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
  const uint64_t a = (static_cast<uint64_t>(r.x) << 21) ^ (r.y >> 11);  // 53 bits
  const uint64_t b = (static_cast<uint64_t>(r.z) << 21) ^ (r.w >> 11);
  const double u1 = (static_cast<double>(a & ((1ull << 53) - 1)) + 1.0) * 0x1.0p-53;  // (0, 1]
  const double u2 = static_cast<double>(b & ((1ull << 53) - 1)) * 0x1.0p-53;          // [0, 1)
  const double rad = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  z0 = rad * c;
  z1 = rad * s;
}

// The nearest FP16 value of z (through FP32; stored as FP64): an N(0,1) sample that the FP16
// sketch packs without rounding.
__device__ __forceinline__ double to_f16_grid(double z) {
  return static_cast<double>(static_cast<float>(static_cast<_Float16>(static_cast<float>(z))));
}

template <bool F16_GRID>
__global__ void omega_philox_kernel(double* __restrict__ out, int64_t count, uint64_t seed) {
  const int64_t npairs = (count + 1) / 2;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npairs;
       p += (int64_t)gridDim.x * blockDim.x) {
    double z0, z1;
    gaussian_pair(seed, static_cast<uint64_t>(p), z0, z1);
    if (F16_GRID) {
      z0 = to_f16_grid(z0);
      z1 = to_f16_grid(z1);
    }
    out[2 * p] = z0;
    if (2 * p + 1 < count) out[2 * p + 1] = z1;
  }
}

// A(i, j) = sum_t U(i, t) sigma(t) V(j, t) + noise * g(i + j * m): the noise of element (i, j)
// is element i + j*m of the Philox stream, so the same A results for both storage orders.
// Threads walk the storage order so that writes are coalesced.
template <typename T>
__global__ void synth_kernel(int64_t m, int64_t n, int64_t r, const double* __restrict__ U,
                             const double* __restrict__ V, const double* __restrict__ sigma,
                             double noise, uint64_t seed, T* __restrict__ A, int64_t lda,
                             bool colmajor) {
  const int64_t total = m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = colmajor ? e % m : e / n;
    const int64_t j = colmajor ? e / m : e % n;
    double acc = 0.0;
    for (int64_t t = 0; t < r; ++t) acc += U[i * r + t] * sigma[t] * V[j * r + t];
    if (noise != 0.0) {
      const int64_t canon = i + j * m;
      double z0, z1;
      gaussian_pair(seed, static_cast<uint64_t>(canon >> 1), z0, z1);
      acc += noise * ((canon & 1) ? z1 : z0);
    }
    A[colmajor ? (i + j * lda) : (i * lda + j)] = static_cast<T>(acc);
  }
}

// A = T + noise * g with T row-major m x n FP64 (the low-rank part from a GEMM), written in A's
// storage order; the noise is the same stream as in synth_kernel.
template <typename T>
__global__ void synth_finish_kernel(int64_t m, int64_t n, const double* __restrict__ Tm,
                                    double noise, uint64_t seed, T* __restrict__ A, int64_t lda,
                                    bool colmajor) {
  const int64_t total = m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = colmajor ? e % m : e / n;
    const int64_t j = colmajor ? e / m : e % n;
    double acc = Tm[i * n + j];
    if (noise != 0.0) {
      const int64_t canon = i + j * m;
      double z0, z1;
      gaussian_pair(seed, static_cast<uint64_t>(canon >> 1), z0, z1);
      acc += noise * ((canon & 1) ? z1 : z0);
    }
    A[colmajor ? (i + j * lda) : (i * lda + j)] = static_cast<T>(acc);
  }
}

// Y (m x nl, row-major FP64) = columns idx[0..nl) of A (element (i, k) at A[i + k*lda] when
// col-major, A[i*lda + k] when row-major).  One element per thread, grid-stride; Y is written in
// order.
template <typename T>
__global__ void gather_columns_kernel(const T* __restrict__ A, int64_t lda, bool colmajor,
                                      int64_t m, const int64_t* __restrict__ idx, int64_t nl,
                                      double* __restrict__ Y) {
  const int64_t total = m * nl;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = e / nl, c = e % nl;
    const int64_t k = idx[c];
    Y[e] = static_cast<double>(colmajor ? A[i + k * lda] : A[i * lda + k]);
  }
}

// A = (I - 2uu') Dt (I - 2vv'), Dt = [diag(sigma); 0] (m x n, n <= m), u (m) and v (n) unit
// vectors: A(i, j) = [i == j] sigma_j - 2 (sigma_i v_i) v_j - 2 u_i (sigma_j u_j) + 4 c u_i v_j
// with sigma_i v_i = 0 for i >= n and c = u(0:n)' diag(sigma) v.  A block writes `lines`
// consecutive storage lines (columns of a column-major A, rows of a row-major one) with its
// threads along each line, so writes are coalesced and no element index is divided; every
// block forms c itself with the same reduction order (256 threads), so all use the same value.
template <typename T>
__global__ void __launch_bounds__(256) synth_reflected_kernel(
    int64_t m, int64_t n, const double* __restrict__ sigma, const double* __restrict__ u,
    const double* __restrict__ v, T* __restrict__ A, int64_t lda, bool colmajor, int64_t lines) {
  __shared__ double red[4];
  double part = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256) part += u[j] * sigma[j] * v[j];
  part = wave_sum(part);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  const double c = (red[0] + red[1]) + (red[2] + red[3]);
  const int64_t nlines = colmajor ? n : m;
  const int64_t l0 = blockIdx.x * lines;
  const int64_t l1 = l0 + lines < nlines ? l0 + lines : nlines;
  for (int64_t line = l0; line < l1; ++line) {
    T* out = A + line * lda;
    if (colmajor) {  // column j = line; i runs along it
      const int64_t j = line;
      const double vj = v[j], bj = sigma[j] * u[j], sj = sigma[j];
      for (int64_t i = threadIdx.x; i < m; i += 256) {
        const double ai = i < n ? sigma[i] * v[i] : 0.0;
        double x = -2.0 * ai * vj - 2.0 * u[i] * bj + 4.0 * c * u[i] * vj;
        if (i == j) x += sj;
        out[i] = static_cast<T>(x);
      }
    } else {  // row i = line; j runs along it
      const int64_t i = line;
      const double ai = i < n ? sigma[i] * v[i] : 0.0, ui = u[i];
      for (int64_t j = threadIdx.x; j < n; j += 256) {
        double x = -2.0 * ai * v[j] - 2.0 * ui * (sigma[j] * u[j]) + 4.0 * c * ui * v[j];
        if (i == j) x += sigma[j];
        out[j] = static_cast<T>(x);
      }
    }
  }
}

__global__ void scale_stream_kernel(double* __restrict__ out, int64_t count, double scale) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] *= scale;
}

}  // namespace

void synth_finish(int64_t m, int64_t n, const double* Tm, double noise, uint64_t seed, void* A,
                  int64_t lda, bool a_f32, bool a_colmajor, hipStream_t st) {
  if (m <= 0 || n <= 0) return;
  int64_t blocks = ceil_div(m * n, 256);
  if (blocks > 65536) blocks = 65536;
  if (a_f32)
    hipLaunchKernelGGL(synth_finish_kernel<float>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       st, m, n, Tm, noise, seed, static_cast<float*>(A), lda, a_colmajor);
  else
    hipLaunchKernelGGL(synth_finish_kernel<double>, dim3(static_cast<unsigned>(blocks)), dim3(256),
                       0, st, m, n, Tm, noise, seed, static_cast<double*>(A), lda, a_colmajor);
  MCNLA_LAUNCHED();
}

void gather_columns(const void* A, int64_t lda, bool a_f32, bool a_colmajor, int64_t m,
                    const int64_t* idx, int64_t nl, double* Y, hipStream_t st) {
  if (m <= 0 || nl <= 0) return;
  int64_t blocks = ceil_div(m * nl, 256);
  if (blocks > 65536) blocks = 65536;
  if (a_f32)
    hipLaunchKernelGGL(gather_columns_kernel<float>, dim3(static_cast<unsigned>(blocks)), dim3(256),
                       0, st, static_cast<const float*>(A), lda, a_colmajor, m, idx, nl, Y);
  else
    hipLaunchKernelGGL(gather_columns_kernel<double>, dim3(static_cast<unsigned>(blocks)), dim3(256),
                       0, st, static_cast<const double*>(A), lda, a_colmajor, m, idx, nl, Y);
  MCNLA_LAUNCHED();
}

void gaussian_fill(double* out, int64_t count, uint64_t seed, double scale, hipStream_t st) {
  omega_philox(out, count, seed, st);
  if (scale != 1.0 && count > 0) {
    int64_t blocks = ceil_div(count, 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(scale_stream_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                       out, count, scale);
    MCNLA_LAUNCHED();
  }
}

void omega_philox(double* out, int64_t count, uint64_t seed, hipStream_t st, bool f16_grid) {
  if (count <= 0) return;
  const int threads = 256;
  int64_t blocks = ceil_div((count + 1) / 2, threads);
  if (blocks > 65536) blocks = 65536;
  if (f16_grid)
    hipLaunchKernelGGL(omega_philox_kernel<true>, dim3(static_cast<unsigned>(blocks)),
                       dim3(threads), 0, st, out, count, seed);
  else
    hipLaunchKernelGGL(omega_philox_kernel<false>, dim3(static_cast<unsigned>(blocks)),
                       dim3(threads), 0, st, out, count, seed);
  MCNLA_LAUNCHED();
}

void synth_lowrank(int64_t m, int64_t n, int64_t r, const double* U, const double* V,
                   const double* sigma, double noise, uint64_t seed, void* A, int64_t lda,
                   bool a_f32, bool a_colmajor, hipStream_t st) {
  if (m <= 0 || n <= 0) return;
  const int threads = 256;
  int64_t blocks = ceil_div(m * n, threads);
  if (blocks > 65536) blocks = 65536;
  if (a_f32) {
    hipLaunchKernelGGL(synth_kernel<float>, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0,
                       st, m, n, r, U, V, sigma, noise, seed, static_cast<float*>(A), lda,
                       a_colmajor);
  } else {
    hipLaunchKernelGGL(synth_kernel<double>, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0,
                       st, m, n, r, U, V, sigma, noise, seed, static_cast<double*>(A), lda,
                       a_colmajor);
  }
  MCNLA_LAUNCHED();
}

void synth_reflected(int64_t m, int64_t n, const double* sigma, const double* u, const double* v,
                     void* A, int64_t lda, bool a_f32, bool a_colmajor, hipStream_t st) {
  if (m <= 0 || n <= 0) return;
  const int64_t nlines = a_colmajor ? n : m;
  const int64_t lines = ceil_div(nlines, 1024);  // about 1024 blocks
  const unsigned blocks = static_cast<unsigned>(ceil_div(nlines, lines));
  if (a_f32)
    hipLaunchKernelGGL(synth_reflected_kernel<float>, dim3(blocks), dim3(256), 0, st, m, n, sigma,
                       u, v, static_cast<float*>(A), lda, a_colmajor, lines);
  else
    hipLaunchKernelGGL(synth_reflected_kernel<double>, dim3(blocks), dim3(256), 0, st, m, n,
                       sigma, u, v, static_cast<double*>(A), lda, a_colmajor, lines);
  MCNLA_LAUNCHED();
}

}  // namespace mcnla
