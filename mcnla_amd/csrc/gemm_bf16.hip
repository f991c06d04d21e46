// Mixed-precision sketch Y = A * Omega on BF16 MFMA with FP32 accumulation (gfx950).
//
// Pass 1 packs A (m x n, FP32/FP64, either order) and Omega (n x Nl, FP64) into K-contiguous,
// zero-padded BF16 panels: Ap[Mp][Kp] and Wp[Np][Kp] (K = n).  With terms == 3 each operand is
// split x = hi + lo (both BF16, lo = bf16(x - hi)) and the GEMM accumulates
// hi*hi + lo*hi + hi*lo, which carries ~16 mantissa bits of each operand.
// Pass 2 is the GEMM: 128 x 128 tiles, 4 waves of 64 x 64, v_mfma_f32_16x16x32_bf16; each lane
// reads 8 consecutive k of a row from LDS (one 16-byte read per fragment).
#include "common.hpp"
#include "kernels.hpp"

namespace mcnla {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128;
constexpr int kThreads = 256;
constexpr int kTile = 64;  // pack transpose tile

__host__ __device__ constexpr int bk_for(int terms) { return terms == 1 ? 64 : 32; }

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xFFFF) ? 0x40 : 0));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// Packs a matrix whose element (r, k) lives at src[r + k*ld] (R_CONTIG) or src[r*ld + k] into
// dst_hi/dst_lo [rows_pad][kpad] (k contiguous), zero outside (rows, cols).
template <typename TS, bool R_CONTIG>
__global__ __launch_bounds__(kThreads) void pack_kernel(const TS* __restrict__ src, int64_t ld,
                                                        int64_t rows, int64_t cols,
                                                        int64_t kpad, uint16_t* __restrict__ hi,
                                                        uint16_t* __restrict__ lo) {
  __shared__ float tile[kTile][kTile + 1];
  const int64_t r0 = blockIdx.y * (int64_t)kTile, k0 = blockIdx.x * (int64_t)kTile;
  for (int e = threadIdx.x; e < kTile * kTile; e += kThreads) {
    int rr, kk;
    if (R_CONTIG) {
      kk = e / kTile;
      rr = e % kTile;
    } else {
      rr = e / kTile;
      kk = e % kTile;
    }
    const int64_t gr = r0 + rr, gk = k0 + kk;
    float v = 0.0f;
    if (gr < rows && gk < cols) v = static_cast<float>(R_CONTIG ? src[gr + gk * ld] : src[gr * ld + gk]);
    tile[rr][kk] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kTile * kTile; e += kThreads) {
    const int rr = e / kTile, kk = e % kTile;
    const int64_t o = (r0 + rr) * kpad + k0 + kk;
    const float v = tile[rr][kk];
    const uint16_t h = f32_to_bf16_rne(v);
    hi[o] = h;
    if (lo) lo[o] = f32_to_bf16_rne(v - bf16_to_f32(h));
  }
}

template <int NOPS, int LOADS>
__device__ __forceinline__ void tile_gload(u32x4 (&ra)[NOPS][LOADS], u32x4 (&rb)[NOPS][LOADS],
                                           const uint16_t* __restrict__ a_hi,
                                           const uint16_t* __restrict__ a_lo,
                                           const uint16_t* __restrict__ b_hi,
                                           const uint16_t* __restrict__ b_lo,
                                           const int64_t (&goff)[LOADS], int64_t k0) {
#pragma unroll
  for (int it = 0; it < LOADS; ++it) {
    ra[0][it] = *reinterpret_cast<const u32x4*>(a_hi + goff[it] + k0);
    rb[0][it] = *reinterpret_cast<const u32x4*>(b_hi + goff[it] + k0);
    if (NOPS == 2) {
      ra[NOPS - 1][it] = *reinterpret_cast<const u32x4*>(a_lo + goff[it] + k0);
      rb[NOPS - 1][it] = *reinterpret_cast<const u32x4*>(b_lo + goff[it] + k0);
    }
  }
}

template <int NOPS, int LOADS, int PANEL>
__device__ __forceinline__ void tile_lstore_impl(uint16_t (*sa)[PANEL], uint16_t (*sb)[PANEL],
                                                 const u32x4 (&ra)[NOPS][LOADS],
                                                 const u32x4 (&rb)[NOPS][LOADS],
                                                 const int (&loff)[LOADS]) {
#pragma unroll
  for (int o = 0; o < NOPS; ++o)
#pragma unroll
    for (int it = 0; it < LOADS; ++it) {
      *reinterpret_cast<u32x4*>(&sa[o][loff[it]]) = ra[o][it];
      *reinterpret_cast<u32x4*>(&sb[o][loff[it]]) = rb[o][it];
    }
}


template <int TERMS, typename TY>
__global__ __launch_bounds__(kThreads) void sketch_bf16_kernel(
    const uint16_t* __restrict__ Ahi, const uint16_t* __restrict__ Alo,
    const uint16_t* __restrict__ Whi, const uint16_t* __restrict__ Wlo, int64_t kpad, int64_t m,
    int64_t nl, TY* __restrict__ Y, int64_t ldy) {
  constexpr int BK = bk_for(TERMS);
  constexpr int NOPS = TERMS == 1 ? 1 : 2;   // hi (+ lo) panels per operand
  constexpr int LDK = BK + 8;                // bf16 per LDS row (16-byte pad)
  constexpr int CH = BK / 8;                 // 16-byte chunks per row
  constexpr int LOADS = kBM * CH / kThreads; // 16-byte loads per thread per panel
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][NOPS][kBM * LDK];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][NOPS][kBN * LDK];

  const int64_t tiles_n = ceil_div(nl, kBN);
  const int64_t tiles_m = ceil_div(m, kBM);
  const int64_t nblocks = tiles_m * tiles_n;
  const int64_t logical = xcd_swizzle(blockIdx.x, nblocks);
  const int64_t tm = logical / tiles_n, tn = logical % tiles_n;
  const int64_t m0 = tm * kBM, n0 = tn * kBN;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int wm0 = (wave / 2) * 64, wn0 = (wave % 2) * 64;

  // Per-thread global offsets of its 16-byte chunks (row r, chunk c) inside a K tile.
  const uint16_t* __restrict__ a_hi = Ahi + m0 * kpad;
  const uint16_t* __restrict__ b_hi = Whi + n0 * kpad;
  const uint16_t* __restrict__ a_lo = (NOPS == 2) ? Alo + m0 * kpad : a_hi;
  const uint16_t* __restrict__ b_lo = (NOPS == 2) ? Wlo + n0 * kpad : b_hi;
  int64_t goff[LOADS];
  int loff[LOADS];
#pragma unroll
  for (int it = 0; it < LOADS; ++it) {
    const int idx = it * kThreads + threadIdx.x;
    const int r = idx / CH, c = idx % CH;
    goff[it] = r * kpad + c * 8;
    loff[it] = r * LDK + c * 8;
  }

  u32x4 ra[NOPS][LOADS], rb[NOPS][LOADS];
  tile_gload<NOPS, LOADS>(ra, rb, a_hi, a_lo, b_hi, b_lo, goff, 0);
  tile_lstore_impl<NOPS, LOADS>(sA[0], sB[0], ra, rb, loff);

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  const int64_t ntk = kpad / BK;
  __syncthreads();
  for (int64_t kt = 0; kt < ntk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < ntk;
    if (more) tile_gload<NOPS, LOADS>(ra, rb, a_hi, a_lo, b_hi, b_lo, goff, (kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 32) {
      const int kof = kk + 8 * (lane / 16);
      bf16x8 af[NOPS][4], bfr[NOPS][4];
#pragma unroll
      for (int o = 0; o < NOPS; ++o)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          af[o][i] = *reinterpret_cast<const bf16x8*>(&sA[cur][o][(wm0 + i * 16 + lane % 16) * LDK + kof]);
          bfr[o][i] = *reinterpret_cast<const bf16x8*>(&sB[cur][o][(wn0 + i * 16 + lane % 16) * LDK + kof]);
        }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfr[0][j], acc[i][j], 0, 0, 0);
          if (TERMS == 3) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[NOPS - 1][i], bfr[0][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfr[NOPS - 1][j], acc[i][j], 0, 0, 0);
          }
        }
    }
    if (more) tile_lstore_impl<NOPS, LOADS>(sA[cur ^ 1], sB[cur ^ 1], ra, rb, loff);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = m0 + wm0 + i * 16 + 4 * (lane / 16) + r;
        const int64_t gj = n0 + wn0 + j * 16 + lane % 16;
        if (gi < m && gj < nl) Y[gi * ldy + gj] = static_cast<TY>(acc[i][j][r]);
      }
}

struct PackDims {
  int64_t mp, np, kp;
};
PackDims pack_dims(const SketchBf16Args& a) {
  const int bk = 64;  // pad K to the larger K tile so both term counts fit
  return PackDims{ceil_div(a.m, kBM) * kBM, ceil_div(a.nl, kBN) * kBN, ceil_div(a.n, bk) * bk};
}

template <typename TS>
void launch_pack(const TS* src, int64_t ld, bool r_contig, int64_t rows, int64_t cols,
                 int64_t rows_pad, int64_t kpad, uint16_t* hi, uint16_t* lo, hipStream_t st) {
  dim3 grid(static_cast<unsigned>(ceil_div(kpad, kTile)), static_cast<unsigned>(ceil_div(rows_pad, kTile)));
  if (r_contig)
    hipLaunchKernelGGL((pack_kernel<TS, true>), grid, dim3(kThreads), 0, st, src, ld, rows, cols,
                       kpad, hi, lo);
  else
    hipLaunchKernelGGL((pack_kernel<TS, false>), grid, dim3(kThreads), 0, st, src, ld, rows, cols,
                       kpad, hi, lo);
  MCNLA_LAUNCHED();
}

}  // namespace

size_t sketch_bf16_workspace(const SketchBf16Args& a) {
  const PackDims d = pack_dims(a);
  const size_t per = (a.terms == 3) ? 2 : 1;
  // Mp and Np are multiples of 128 and Kp of 64, so each panel is a multiple of 16 KiB.
  return per * static_cast<size_t>(d.kp) * static_cast<size_t>(d.mp + d.np) * sizeof(uint16_t);
}

void sketch_bf16(const SketchBf16Args& a, void* workspace, size_t workspace_bytes, hipStream_t st,
                 hipEvent_t gemm_begin, hipEvent_t gemm_end) {
  MCNLA_REQUIRE(a.terms == 1 || a.terms == 3, MCNLA_ERR_INVALID, "terms must be 1 or 3");
  MCNLA_REQUIRE(workspace_bytes >= sketch_bf16_workspace(a), MCNLA_ERR_INVALID,
                "bf16 sketch workspace too small");
  if (a.m == 0 || a.nl == 0) return;
  const PackDims d = pack_dims(a);
  uint16_t* base = static_cast<uint16_t*>(workspace);
  uint16_t* Ahi = base;
  uint16_t* Whi = Ahi + d.mp * d.kp;
  uint16_t* Alo = nullptr;
  uint16_t* Wlo = nullptr;
  if (a.terms == 3) {
    Alo = Whi + d.np * d.kp;
    Wlo = Alo + d.mp * d.kp;
  }
  // A: element (i, k) at A[i + k*lda] (col-major) or A[i*lda + k] (row-major).
  if (a.a_f32)
    launch_pack(static_cast<const float*>(a.A), a.lda, a.a_colmajor, a.m, a.n, d.mp, d.kp, Ahi, Alo, st);
  else
    launch_pack(static_cast<const double*>(a.A), a.lda, a.a_colmajor, a.m, a.n, d.mp, d.kp, Ahi, Alo, st);
  // Omega^T: element (j, k) = Omega[k*nl + j], i.e. row index j contiguous.
  launch_pack(a.omega, a.nl, true, a.nl, a.n, d.np, d.kp, Whi, Wlo, st);
  const int64_t nblocks = ceil_div(a.m, kBM) * ceil_div(a.nl, kBN);
  if (gemm_begin) MCNLA_HIP(hipEventRecord(gemm_begin, st));
  const dim3 grid(static_cast<unsigned>(nblocks));
  if (a.terms == 1 && a.Yd)
    hipLaunchKernelGGL((sketch_bf16_kernel<1, double>), grid, dim3(kThreads), 0, st, Ahi, Alo, Whi,
                       Wlo, d.kp, a.m, a.nl, a.Yd, a.ldy);
  else if (a.terms == 1)
    hipLaunchKernelGGL((sketch_bf16_kernel<1, float>), grid, dim3(kThreads), 0, st, Ahi, Alo, Whi,
                       Wlo, d.kp, a.m, a.nl, a.Y, a.ldy);
  else if (a.Yd)
    hipLaunchKernelGGL((sketch_bf16_kernel<3, double>), grid, dim3(kThreads), 0, st, Ahi, Alo, Whi,
                       Wlo, d.kp, a.m, a.nl, a.Yd, a.ldy);
  else
    hipLaunchKernelGGL((sketch_bf16_kernel<3, float>), grid, dim3(kThreads), 0, st, Ahi, Alo, Whi,
                       Wlo, d.kp, a.m, a.nl, a.Y, a.ldy);
  MCNLA_LAUNCHED();
  if (gemm_end) MCNLA_HIP(hipEventRecord(gemm_end, st));
}

}  // namespace mcnla
