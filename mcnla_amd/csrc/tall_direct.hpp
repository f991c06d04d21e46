// The tall-skinny FP64-MFMA product tile, shared by the GEMM dispatch (gemm_f64.hip) and the
// integrators' grouped Bs products (integrate.hip), so that both run the same arithmetic.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace mcnla {

template <typename TS, int VEC>
struct VecOf;
template <>
struct VecOf<double, 1> { typedef double type; };
template <>
struct VecOf<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };
template <>
struct VecOf<float, 1> { typedef float type; };
template <>
struct VecOf<float, 2> { typedef float type __attribute__((ext_vector_type(2))); };
template <>
struct VecOf<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };

// Tall products with the A operand read from HBM straight into MFMA registers (no LDS slab for
// A).  With VE = 16 / sizeof(TA) elements per 16-byte load, lane L of a wave owns row L % 16 of
// each of its two 16-row groups, and load q of a 32-deep step gives it
// k = 4 VE q + VE (L / 16) + e (e < VE): the 4 lanes of a row read 64 contiguous bytes per
// instruction.  MFMA j contracts over k = kof(j) = 4 VE (j / VE) + VE (L / 16) + j % VE, the same
// k for the B rows its lanes read from LDS (any k order gives the sum; this one pairs A and B
// lanes).  Only the 32 x LP slab of B goes through LDS, so more workgroups fit on a CU and more
// of A is in flight than when A is staged too; the next step's A (other register set) and B load
// during the current MFMAs.  The B slab's row stride LP + 4 puts the 4 k-rows a wave reads at
// once on the 64 banks at most twice each.
//
// TM is the arithmetic: double (v_mfma_f64_16x16x4f64; an FP32 A is widened on load) or float
// (FP32 A, B rounded to FP32, v_mfma_f32_16x16x4f32 with FP32 sums within one split's K range;
// the split-K partials are summed in FP64).  The GEMM dispatch uses it for FP64 A; for an FP32 A
// the LDS-staged tall_kernel above measured as fast or faster (mixed former Z: 1.25 against
// 1.31 ms; 500000 x 2000 FP32 A times X: equal).
//
// Computes workgroup `block` of the product (tile block % tiles_m, split block / tiles_m) on
// 256 threads; the caller launches ceil(M / 128) * splits of them.
template <typename TA, typename TM, int LP, bool VEC>
__device__ __forceinline__ void tall_direct_tile(const GemmF64Args& p, int64_t block) {
  constexpr int BM = 128, BK = 32;
  constexpr int SB = LP + 4;
  constexpr int NB = BK * LP / 256;  // B elements per thread and step
  constexpr int TJ = LP / 16;
  constexpr int VE = 16 / static_cast<int>(sizeof(TA));
  constexpr bool F32 = sizeof(TM) == 4;
  static_assert(!F32 || sizeof(TA) == 4, "FP32 MFMA needs an FP32 A");
  typedef typename VecOf<TA, VE>::type V;
  typedef typename std::conditional<F32, f4, d4>::type Acc;
  __shared__ TM sB[BK * SB];
  const int64_t tiles_m = ceil_div(p.M, BM);
  const int64_t tile = block % tiles_m, split = block / tiles_m;
  const int64_t i0 = tile * BM;
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const int fr = lane / 16, fc = lane % 16;
  const TA* A = static_cast<const TA*>(p.A);
  const double* B = static_cast<const double*>(p.B);
  const TA* arow[2];
  bool rok[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int64_t row = i0 + wave * 32 + a * 16 + fc;
    rok[a] = row < p.M;
    arow[a] = A + (rok[a] ? row : 0) * p.lda;
  }
  auto kof = [&](int j) { return 4 * VE * (j / VE) + VE * fr + (j % VE); };
  auto load_a = [&](int64_t k0, TM (&v)[2][8]) {
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      if (VEC && rok[a] && k0 + BK <= kend) {
#pragma unroll
        for (int q = 0; q < 8 / VE; ++q) {
          const V x = *reinterpret_cast<const V*>(arow[a] + k0 + 4 * VE * q + VE * fr);
#pragma unroll
          for (int e = 0; e < VE; ++e) v[a][VE * q + e] = static_cast<TM>(x[e]);
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          v[a][e] = (rok[a] && k0 + kof(e) < kend) ? static_cast<TM>(arow[a][k0 + kof(e)]) : TM(0);
      }
    }
  };
  TM vb[NB];
  auto load_b = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = e * 256 + tid, kb = idx / LP, jb = idx % LP;
      vb[e] = (jb < p.N && k0 + kb < kend) ? static_cast<TM>(B[(k0 + kb) * p.ldb + jb]) : TM(0);
    }
  };
  Acc acc[2][TJ];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b) acc[a][b] = Acc{0, 0, 0, 0};
  auto step = [&](int64_t k0, const TM (&cur)[2][8], TM (&nxt)[2][8]) {
    __syncthreads();  // the previous step's reads of sB are done
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = e * 256 + tid;
      sB[(idx / LP) * SB + idx % LP] = vb[e];
    }
    __syncthreads();
    if (k0 + BK < kend) {
      load_a(k0 + BK, nxt);
      load_b(k0 + BK);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      TM bf[TJ];
#pragma unroll
      for (int b = 0; b < TJ; ++b) bf[b] = sB[kof(j) * SB + b * 16 + fc];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b) {
          if constexpr (F32)
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[a][j], bf[b], acc[a][b], 0, 0, 0);
          else
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur[a][j], bf[b], acc[a][b], 0, 0, 0);
        }
    }
  };
  TM va0[2][8], va1[2][8];
  if (kbeg < kend) {
    load_a(kbeg, va0);
    load_b(kbeg);
  }
  for (int64_t k0 = kbeg; k0 < kend; k0 += 2 * BK) {
    step(k0, va0, va1);
    if (k0 + BK < kend) step(k0 + BK, va1, va0);
  }
  // D element r of a 16 x 16 tile sits at row lane/16 + 4r (FP64) or 4 (lane/16) + r (FP32),
  // column lane%16
  const bool to_partial = p.partial != nullptr;
  double* P = to_partial ? p.partial + split * p.M * p.N : nullptr;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = i0 + wave * 32 + a * 16 + (F32 ? 4 * fr + r : fr + 4 * r);
        const int64_t gj = b * 16 + fc;
        if (gi >= p.M || gj >= p.N) continue;
        const double v = static_cast<double>(acc[a][b][r]);
        if (to_partial) {
          P[gi * p.N + gj] = v;
        } else {
          double* c = p.C + gi * p.ldc + gj;
          *c = (p.beta == 0.0) ? p.alpha * v : p.alpha * v + p.beta * *c;
        }
      }
}

}  // namespace mcnla
