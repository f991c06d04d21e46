// Mixed-precision sketch Y = A * Omega on BF16 MFMA with FP32 accumulation (gfx950).
//
// One term (MIXED_BF16): A (m x n, FP32/FP64, either order) and Omega (n x Nl, FP64) are rounded
// to BF16 and packed into swizzled 256 x 32 blocks; the GEMM streams the blocks into LDS with
// direct-to-LDS loads (256 x 256 tiles, 4 waves of 128 x 128).
// Three terms (MIXED_BF16X3): each operand is split x = hi + lo (both BF16, lo = bf16(x - hi))
// into K-contiguous zero-padded panels Ap[Mp][Kp], Wp[Np][Kp], and the GEMM (128 x 128 tiles,
// 4 waves of 64 x 64) accumulates hi*hi + lo*hi + hi*lo, which carries ~16 mantissa bits of each
// operand.  Both use v_mfma_f32_16x16x32_bf16 with each lane reading 8 consecutive k of a row
// from LDS (one 16-byte read per fragment).
#include "common.hpp"
#include "kernels.hpp"

namespace mcnla {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// acc += a * b on the matrix cores with the accumulator tied to the destination.  Through the
// builtin, the register allocator may give each MFMA a fresh destination (the untied form's
// destination is early-clobber) and then copy accumulators back: with 256 live accumulator
// registers that is more AGPR moves than MFMAs, and every move waits for the MFMA it reads.
// Inline asm is outside the compiler's hazard tracking, so a caller keeps consecutive MFMAs on
// one accumulator far apart and pads with s_nop before the first MFMA and before VALU reads of
// the results.
__device__ __forceinline__ void mfma_bf16_tied(f4& acc, const bf16x8& a, const bf16x8& b) {
  asm volatile("v_mfma_f32_16x16x32_bf16 %0, %1, %2, %0" : "+a"(acc) : "v"(a), "v"(b));
}

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

// ------------------------------------------------------------------------------------------------
// One-term path.  Both operands are packed into 256-row x 32-k BF16 blocks of 16 KiB, laid out
// [row tile][k block][256][32] so that every pipeline stage of the GEMM is one contiguous block
// per operand.  Inside a block the 16-byte chunk c = k/8 of row r is stored at chunk position
// c ^ ((r / 4) % 4): a linear copy into LDS then gives fragment reads (16 rows x one chunk per
// 16 lanes) that hit 16 distinct 4-bank groups.  The k blocks are padded with zero blocks to a
// multiple of kUnroll.
constexpr int kBR = 256;       // rows per block
constexpr int kBKB = 32;       // k per block
constexpr int kBlk = kBR * kBKB;  // elements per block
constexpr int kStages = 5;   // LDS ring depth of the GEMM (5 x 32 KiB)
constexpr int kUnroll = 10;  // steps per loop body: ring slots x fragment buffers

__host__ __device__ constexpr int blk_off(int r, int k) {
  return r * kBKB + 8 * ((k >> 3) ^ ((r >> 2) & 3)) + (k & 7);
}

int64_t blk_nkt(int64_t n) { return ceil_div(ceil_div(n, kBKB), kUnroll) * kUnroll; }

// src element (r, k) at src[r + k*ld] (R_CONTIG) or src[r*ld + k]; one workgroup per block.
// VEC4: FP32 R_CONTIG source with ld % 4 == 0 and a 16-byte aligned base (float4 loads).
template <typename TS, bool R_CONTIG, bool VEC4>
__global__ __launch_bounds__(256) void pack_blocks_kernel(const TS* __restrict__ src, int64_t ld,
                                                          int64_t rows, int64_t cols, int64_t nkt,
                                                          uint16_t* __restrict__ dst) {
  constexpr int LDT = kBR + 4;  // floats per LDS row (rows stay 16-byte aligned)
  __shared__ __attribute__((aligned(16))) float tile[kBKB][LDT];
  const int64_t tr = blockIdx.y, kt = blockIdx.x;
  const int64_t r0 = tr * kBR, k0 = kt * kBKB;
  const int t = threadIdx.x;
  if constexpr (VEC4 && R_CONTIG) {
    const bool full = r0 + kBR <= rows;
#pragma unroll
    for (int j = 0; j < 8; ++j) {  // a wave reads one 1 KiB column segment per instruction
      const int k = j * 4 + t / 64, r = 4 * (t % 64);
      const int64_t gk = k0 + k;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gk < cols) {
        const float* col = reinterpret_cast<const float*>(src) + gk * ld + r0 + r;
        if (full) {
          v = *reinterpret_cast<const float4*>(col);
        } else {
          v.x = (r0 + r < rows) ? col[0] : 0.f;
          v.y = (r0 + r + 1 < rows) ? col[1] : 0.f;
          v.z = (r0 + r + 2 < rows) ? col[2] : 0.f;
          v.w = (r0 + r + 3 < rows) ? col[3] : 0.f;
        }
      }
      *reinterpret_cast<float4*>(&tile[k][r]) = v;
    }
  } else {
    for (int e = t; e < kBR * kBKB; e += 256) {
      int r, k;
      if (R_CONTIG) {
        k = e / kBR;
        r = e % kBR;
      } else {
        r = e / kBKB;
        k = e % kBKB;
      }
      const int64_t gr = r0 + r, gk = k0 + k;
      tile[k][r] = (gr < rows && gk < cols)
                       ? static_cast<float>(R_CONTIG ? src[gr + gk * ld] : src[gr * ld + gk])
                       : 0.0f;
    }
  }
  __syncthreads();
  // every store instruction writes 1 KiB contiguous: chunk g = q*256 + t is (row g/4, slot g%4)
  uint16_t* blk = dst + (tr * nkt + kt) * kBlk;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int g = q * 256 + t;
    const int r = g >> 2;
    const int c = (g & 3) ^ ((r >> 2) & 3);
    bf16x8 h;
#pragma unroll
    for (int i = 0; i < 8; ++i) h[i] = static_cast<__bf16>(tile[8 * c + i][r]);  // RNE
    *reinterpret_cast<u32x4*>(blk + g * 8) = __builtin_bit_cast(u32x4, h);
  }
}

typedef __attribute__((address_space(3))) void lds_void;

// vmcnt <= VM and lgkmcnt == 0 (every LDS read issued so far has completed).  Being a builtin,
// the compiler's wait insertion sees it, so a later wait need not cover those reads again.
template <int VM>
__device__ __forceinline__ void wait_vmcnt_lgkm0() {
  __builtin_amdgcn_s_waitcnt((VM & 0xF) | (0x7 << 4) | (((VM >> 4) & 3) << 14));
}

template <int VM>
__device__ __forceinline__ void wait_vmcnt() {
  // gfx9 s_waitcnt fields: vmcnt[3:0], expcnt[6:4], lgkmcnt[11:8], vmcnt high bits [15:14]
  __builtin_amdgcn_s_waitcnt((VM & 0xF) | (0x7 << 4) | (0xF << 8) | (((VM >> 4) & 3) << 14));
}

// Y (m x nl) = A * Omega from blocks Ab[tm][kt] and Wb[tn][kt]: 256 x 256 tiles, 4 waves of
// 128 x 128 (v_mfma_f32_16x16x32_bf16, 256 accumulators per lane).  Blocks stream from HBM into
// LDS with global_load_lds_dwordx4 through a four-stage ring (three stages in flight).  The ring
// buffers are distinct LDS arrays with static indices and the loop body has no branches (the
// stage three ahead is issued every step, clamped to the last block; nkt is a multiple of 4), so
// the compiler keeps per-buffer wait tracking and the only waits are the explicit ones.
template <typename TY>
__global__ __launch_bounds__(256) void sketch_blk_kernel(const uint16_t* __restrict__ Ab,
                                                        const uint16_t* __restrict__ Wb,
                                                        int64_t nkt, int64_t m, int64_t nl,
                                                        TY* __restrict__ Y, int64_t ldy) {
  constexpr int L = 8;  // DMA instructions per wave and stage (4 for A, 4 for B)
  static_assert(kStages == 5 && kUnroll == 10, "the loop body below is written for 5 slots");
  __shared__ __attribute__((aligned(16))) uint16_t S0[2 * kBlk], S1[2 * kBlk], S2[2 * kBlk],
      S3[2 * kBlk], S4[2 * kBlk];
  const int64_t tiles_n = ceil_div(nl, kBR), tiles_m = ceil_div(m, kBR);
  const int64_t logical = xcd_swizzle(blockIdx.x, tiles_m * tiles_n);
  const int64_t tm = logical / tiles_n, tn = logical % tiles_n;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int wm0 = (wave / 2) * 128, wn0 = (wave % 2) * 128;
  const int kg = lane / 16;
  const uint16_t* abase = Ab + tm * nkt * kBlk;
  const uint16_t* bbase = Wb + tn * nkt * kBlk;
  auto issue = [&](int64_t kt, uint16_t* sd) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int off = (q * 4 + wave) * 512;  // 1 KiB per wave instruction
      __builtin_amdgcn_global_load_lds(abase + kt * kBlk + off + lane * 8, (lds_void*)&sd[off], 16,
                                       0, 0);
      __builtin_amdgcn_global_load_lds(bbase + kt * kBlk + off + lane * 8,
                                       (lds_void*)&sd[kBlk + off], 16, 0, 0);
    }
  };
  f4 acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // Fragments are double-buffered: step kt reads stage kt + 1's fragments from LDS while its
  // MFMAs consume stage kt's, which were read one step earlier.
  bf16x8 fa0[8], fb0[8], fa1[8], fb1[8];
  auto frags = [&](const uint16_t* ss, bf16x8* af, bf16x8* bfr) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      af[i] = *reinterpret_cast<const bf16x8*>(&ss[blk_off(wm0 + i * 16 + lane % 16, 8 * kg)]);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      bfr[j] = *reinterpret_cast<const bf16x8*>(&ss[kBlk + blk_off(wn0 + j * 16 + lane % 16, 8 * kg)]);
  };
  // Step kt: stage kt + 1 must have landed (stages up to kt + 3 are issued; two may still be in
  // flight); the barrier makes that true in every wave.  The stage issued here, kt + 4 (clamped
  // to the last block), overwrites stage kt - 1's slot, whose fragments every wave consumed in
  // step kt - 1, before this barrier.  The last step's look-ahead reads a slot that holds the
  // last block again and discards it.
  auto step = [&](int64_t kt, const bf16x8* caf, const bf16x8* cbf, const uint16_t* snext,
                  bf16x8* naf, bf16x8* nbf, uint16_t* sn) {
    asm volatile("" ::: "memory");
    wait_vmcnt_lgkm0<2 * L>();  // and this step's fragments (read last step) are in registers
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    issue(kt + 4 < nkt ? kt + 4 : nkt - 1, sn);
    frags(snext, naf, nbf);
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) mfma_bf16_tied(acc[i][j], caf[i], cbf[j]);
  };
  issue(0, S0);
  issue(1, S1);
  issue(2, S2);
  issue(3, S3);
  asm volatile("" ::: "memory");
  wait_vmcnt<3 * L>();  // stage 0 landed in every wave
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  frags(S0, fa0, fb0);
  asm volatile("s_nop 7");  // zeroed accumulators (VALU writes) before their first MFMA read
  for (int64_t kt = 0; kt < nkt; kt += kUnroll) {  // stage s sits in slot s % 5
    step(kt, fa0, fb0, S1, fa1, fb1, S4);
    step(kt + 1, fa1, fb1, S2, fa0, fb0, S0);
    step(kt + 2, fa0, fb0, S3, fa1, fb1, S1);
    step(kt + 3, fa1, fb1, S4, fa0, fb0, S2);
    step(kt + 4, fa0, fb0, S0, fa1, fb1, S3);
    step(kt + 5, fa1, fb1, S1, fa0, fb0, S4);
    step(kt + 6, fa0, fb0, S2, fa1, fb1, S0);
    step(kt + 7, fa1, fb1, S3, fa0, fb0, S1);
    step(kt + 8, fa0, fb0, S4, fa1, fb1, S2);
    step(kt + 9, fa1, fb1, S0, fa0, fb0, S3);
  }
  asm volatile("s_nop 7\n s_nop 7\n s_nop 7\n s_nop 7");  // last MFMAs drain before VALU reads
  wait_vmcnt<0>();  // no DMA may still target LDS when the workgroup ends
  const int64_t m0 = tm * kBR, n0 = tn * kBR;
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = m0 + wm0 + i * 16 + 4 * (lane / 16) + r;
        const int64_t gj = n0 + wn0 + j * 16 + lane % 16;
        if (gi < m && gj < nl) Y[gi * ldy + gj] = static_cast<TY>(acc[i][j][r]);
      }
}

struct BlkDims {
  int64_t tm, tn, nkt;
};
BlkDims blk_dims(const SketchBf16Args& a) {
  return BlkDims{ceil_div(a.m, kBR), ceil_div(a.nl, kBR), blk_nkt(a.n)};
}

template <typename TS>
void launch_pack_blocks(const TS* src, int64_t ld, bool r_contig, int64_t rows, int64_t cols,
                        int64_t tiles, int64_t nkt, uint16_t* dst, hipStream_t st) {
  const dim3 grid(static_cast<unsigned>(nkt), static_cast<unsigned>(tiles));
  const bool vec4 = r_contig && ld % 4 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0;
  if constexpr (sizeof(TS) == 4) {
    if (vec4) {
      hipLaunchKernelGGL((pack_blocks_kernel<TS, true, true>), grid, dim3(256), 0, st, src, ld,
                         rows, cols, nkt, dst);
      MCNLA_LAUNCHED();
      return;
    }
  }
  if (r_contig)
    hipLaunchKernelGGL((pack_blocks_kernel<TS, true, false>), grid, dim3(256), 0, st, src, ld, rows,
                       cols, nkt, dst);
  else
    hipLaunchKernelGGL((pack_blocks_kernel<TS, false, false>), grid, dim3(256), 0, st, src, ld,
                       rows, cols, nkt, dst);
  MCNLA_LAUNCHED();
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
  if (a.terms == 1) {
    const BlkDims d = blk_dims(a);
    return static_cast<size_t>(d.tm + d.tn) * static_cast<size_t>(d.nkt) * kBlk * sizeof(uint16_t);
  }
  const PackDims d = pack_dims(a);
  // Mp and Np are multiples of 128 and Kp of 64, so each panel is a multiple of 16 KiB.
  return 2 * static_cast<size_t>(d.kp) * static_cast<size_t>(d.mp + d.np) * sizeof(uint16_t);
}

void sketch_bf16(const SketchBf16Args& a, void* workspace, size_t workspace_bytes, hipStream_t st,
                 hipEvent_t gemm_begin, hipEvent_t gemm_end) {
  MCNLA_REQUIRE(a.terms == 1 || a.terms == 3, MCNLA_ERR_INVALID, "terms must be 1 or 3");
  MCNLA_REQUIRE(workspace_bytes >= sketch_bf16_workspace(a), MCNLA_ERR_INVALID,
                "bf16 sketch workspace too small");
  if (a.m == 0 || a.nl == 0) return;
  if (a.terms == 1) {
    const BlkDims d = blk_dims(a);
    uint16_t* Ab = static_cast<uint16_t*>(workspace);
    uint16_t* Wb = Ab + d.tm * d.nkt * kBlk;
    // A: element (i, k) at A[i + k*lda] (col-major) or A[i*lda + k] (row-major).
    if (a.a_f32)
      launch_pack_blocks(static_cast<const float*>(a.A), a.lda, a.a_colmajor, a.m, a.n, d.tm, d.nkt,
                         Ab, st);
    else
      launch_pack_blocks(static_cast<const double*>(a.A), a.lda, a.a_colmajor, a.m, a.n, d.tm, d.nkt,
                         Ab, st);
    // Omega^T: element (j, k) = Omega[k*nl + j], i.e. row index j contiguous.
    launch_pack_blocks(a.omega, a.nl, true, a.nl, a.n, d.tn, d.nkt, Wb, st);
    const dim3 grid(static_cast<unsigned>(d.tm * d.tn));
    if (gemm_begin) MCNLA_HIP(hipEventRecord(gemm_begin, st));
    if (a.Yd)
      hipLaunchKernelGGL((sketch_blk_kernel<double>), grid, dim3(256), 0, st, Ab, Wb, d.nkt, a.m,
                         a.nl, a.Yd, a.ldy);
    else
      hipLaunchKernelGGL((sketch_blk_kernel<float>), grid, dim3(256), 0, st, Ab, Wb, d.nkt, a.m,
                         a.nl, a.Y, a.ldy);
    MCNLA_LAUNCHED();
    if (gemm_end) MCNLA_HIP(hipEventRecord(gemm_end, st));
    return;
  }
  const PackDims d = pack_dims(a);
  uint16_t* base = static_cast<uint16_t*>(workspace);
  uint16_t* Ahi = base;
  uint16_t* Whi = Ahi + d.mp * d.kp;
  uint16_t* Alo = Whi + d.np * d.kp;
  uint16_t* Wlo = Alo + d.mp * d.kp;
  if (a.a_f32)
    launch_pack(static_cast<const float*>(a.A), a.lda, a.a_colmajor, a.m, a.n, d.mp, d.kp, Ahi, Alo, st);
  else
    launch_pack(static_cast<const double*>(a.A), a.lda, a.a_colmajor, a.m, a.n, d.mp, d.kp, Ahi, Alo, st);
  launch_pack(a.omega, a.nl, true, a.nl, a.n, d.np, d.kp, Whi, Wlo, st);
  const int64_t nblocks = ceil_div(a.m, kBM) * ceil_div(a.nl, kBN);
  if (gemm_begin) MCNLA_HIP(hipEventRecord(gemm_begin, st));
  const dim3 grid(static_cast<unsigned>(nblocks));
  if (a.Yd)
    hipLaunchKernelGGL((sketch_bf16_kernel<3, double>), grid, dim3(kThreads), 0, st, Ahi, Alo, Whi,
                       Wlo, d.kp, a.m, a.nl, a.Yd, a.ldy);
  else
    hipLaunchKernelGGL((sketch_bf16_kernel<3, float>), grid, dim3(kThreads), 0, st, Ahi, Alo, Whi,
                       Wlo, d.kp, a.m, a.nl, a.Y, a.ldy);
  MCNLA_LAUNCHED();
  if (gemm_end) MCNLA_HIP(hipEventRecord(gemm_end, st));
}

}  // namespace mcnla
