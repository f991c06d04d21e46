// Mixed-precision sketch Y = A * Omega on 16-bit MFMA with FP32 accumulation (gfx950).
//
// A (m x n, FP32/FP64, either order) and Omega (n x Nl, FP64) are rounded to BF16 or FP16 and
// packed into swizzled 256 x 32 blocks; one GEMM streams the blocks into LDS with direct-to-LDS
// loads (256 x 256 tiles, 4 waves of 128 x 128, v_mfma_f32_16x16x32_bf16 / _f16).
//
// The more accurate forms add products along K rather than adding kernels: an operand x is split
// into x = hi + lo (lo = round(x - hi)), and the packed K range is a sequence of segments whose
// products sum in the same FP32 accumulators:
//   1 term:  [A] [Omega]
//   2 terms: [A_hi, A_lo] [Omega; Omega]            (A carries about twice the mantissa bits;
//                                                     exact in Omega when Omega is representable
//                                                     in the 16-bit format)
//   3 terms: [A_hi, A_lo, A_hi] [Omega_hi; Omega_hi; Omega_lo]
// FP16 has a 5-bit exponent, so every row i of A is scaled by 2^-e_i before rounding, with e_i
// chosen from the row's largest magnitude so that it lands in [2^13, 2^14); the scale is a power
// of two (exact), and the epilogue multiplies row i of Y by 2^e_i.  A row scale commutes with the
// sum over K, which a per-column scale of A would not.  BF16 shares FP32's exponent range and is
// not scaled.
#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace mcnla {

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kFmtBf16 = 0, kFmtF16 = 1;

// acc += a * b on the matrix cores with the accumulator tied to the destination.  Through the
// builtin, the register allocator may give each MFMA a fresh destination (the untied form's
// destination is early-clobber) and then copy accumulators back: with 256 live accumulator
// registers that is more AGPR moves than MFMAs, and every move waits for the MFMA it reads.
// Inline asm is outside the compiler's hazard tracking, so a caller keeps consecutive MFMAs on
// one accumulator far apart and pads with s_nop before the first MFMA and before VALU reads of
// the results.  a and b hold 8 packed 16-bit values each.
template <int FMT>
__device__ __forceinline__ void mfma_tied(f4& acc, const u32x4& a, const u32x4& b) {
  if constexpr (FMT == kFmtBf16)
    asm volatile("v_mfma_f32_16x16x32_bf16 %0, %1, %2, %0" : "+a"(acc) : "v"(a), "v"(b));
  else
    asm volatile("v_mfma_f32_16x16x32_f16 %0, %1, %2, %0" : "+a"(acc) : "v"(a), "v"(b));
}

// Round to nearest even into the 16-bit format, and back.
template <int FMT>
__device__ __forceinline__ uint16_t to16(float x) {
  if constexpr (FMT == kFmtBf16)
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x));
  else
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(x));
}
template <int FMT>
__device__ __forceinline__ float from16(uint16_t h) {
  if constexpr (FMT == kFmtBf16)
    return static_cast<float>(__builtin_bit_cast(__bf16, h));
  else
    return static_cast<float>(__builtin_bit_cast(_Float16, h));
}

// The exponent e of a row whose largest magnitude has IEEE bits `maxbits`: |a| 2^-e <= 2^14 for
// every entry, the largest in [2^13, 2^14).  Zero rows, and rows holding Inf, keep e = 0 (their
// Inf and NaN values reach Y unchanged).
__device__ __forceinline__ int row_exponent(uint32_t maxbits) {
  const float mx = __uint_as_float(maxbits);
  if (!(mx > 0.f) || !(mx <= 3.4028235e38f)) return 0;
  int e2 = 0;
  (void)frexpf(mx, &e2);  // mx = f 2^e2, f in [0.5, 1)
  return e2 - 14;
}

// rowmax[i] = bits of max_k |A(i, k)| (FP32).  Column-major A: thread per row, K split over
// grid.y, combined with atomicMax on the bit patterns (order-independent; the bits of a
// non-negative float order like the values).  rowmax is zeroed first.
template <typename TS>
__global__ __launch_bounds__(256) void rowmax_colmajor_kernel(const TS* __restrict__ A,
                                                              int64_t lda, int64_t m, int64_t n,
                                                              int64_t kchunk,
                                                              uint32_t* __restrict__ rowmax) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= m) return;
  const int64_t k0 = blockIdx.y * kchunk;
  const int64_t k1 = k0 + kchunk < n ? k0 + kchunk : n;
  float mx = 0.f;
#pragma unroll 8
  for (int64_t k = k0; k < k1; ++k) mx = fmaxf(mx, fabsf(static_cast<float>(A[i + k * lda])));
  atomicMax(&rowmax[i], __float_as_uint(mx));
}

// Row-major A: one wave per row (4 rows per workgroup), written directly.
template <typename TS>
__global__ __launch_bounds__(256) void rowmax_rowmajor_kernel(const TS* __restrict__ A,
                                                              int64_t lda, int64_t m, int64_t n,
                                                              uint32_t* __restrict__ rowmax) {
  const int64_t i = blockIdx.x * 4ll + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (i >= m) return;
  const TS* row = A + i * lda;
  float mx = 0.f;
  for (int64_t k = lane; k < n; k += kWave) mx = fmaxf(mx, fabsf(static_cast<float>(row[k])));
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
  if (lane == 0) rowmax[i] = __float_as_uint(mx);
}

// Operands are packed into 256-row x 32-k blocks of 16 KiB, laid out
// [row tile][segment][k block][256][32] so that every pipeline stage of the GEMM is one contiguous
// block per operand.  Inside a block the 16-byte chunk c = k/8 of row r is stored at chunk
// position c ^ ((r / 4) % 4): a linear copy into LDS then gives fragment reads (16 rows x one
// chunk per 16 lanes) that hit 16 distinct 4-bank groups.  Each segment's k blocks are padded
// with zero blocks to a multiple of kUnroll.
constexpr int kBR = 256;       // rows per block
constexpr int kBKB = 32;       // k per block
constexpr int kBlk = kBR * kBKB;  // elements per block
constexpr int kStages = 5;   // LDS ring depth of the GEMM (5 x 32 KiB)
constexpr int kUnroll = 10;  // steps per loop body: ring slots x fragment buffers

__host__ __device__ constexpr int blk_off(int r, int k) {
  return r * kBKB + 8 * ((k >> 3) ^ ((r >> 2) & 3)) + (k & 7);
}

int64_t blk_nkt(int64_t n) { return ceil_div(ceil_div(n, kBKB), kUnroll) * kUnroll; }

// Packs src element (r, k) (at src[r + k*ld] when R_CONTIG, else src[r*ld + k]) into segment
// `seg_hi` of `nseg` (its rounded value) and, when seg_lo >= 0, into segment seg_lo (the rounded
// remainder).  rowmax (FP16): row r is scaled by 2^-e_r first.  One workgroup per block.
// VEC4: FP32 R_CONTIG source with ld % 4 == 0 and a 16-byte aligned base (float4 loads).
template <typename TS, bool R_CONTIG, bool VEC4, int FMT>
__global__ __launch_bounds__(256) void pack_blocks_kernel(const TS* __restrict__ src, int64_t ld,
                                                          int64_t rows, int64_t cols, int64_t nkt,
                                                          int nseg, int seg_hi, int seg_lo,
                                                          const uint32_t* __restrict__ rowmax,
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
  uint16_t* hi = dst + ((tr * nseg + seg_hi) * nkt + kt) * kBlk;
  uint16_t* lo = dst + ((tr * nseg + (seg_lo < 0 ? 0 : seg_lo)) * nkt + kt) * kBlk;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int g = q * 256 + t;
    const int r = g >> 2;
    const int c = (g & 3) ^ ((r >> 2) & 3);
    const int e = (rowmax && r0 + r < rows) ? row_exponent(rowmax[r0 + r]) : 0;
    uint16_t h[8], l[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const float v = e ? ldexpf(tile[8 * c + i][r], -e) : tile[8 * c + i][r];
      h[i] = to16<FMT>(v);
      l[i] = to16<FMT>(v - from16<FMT>(h[i]));  // exact difference in FP32
    }
    u32x4 ph, pl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ph[i] = static_cast<uint32_t>(h[2 * i]) | (static_cast<uint32_t>(h[2 * i + 1]) << 16);
      pl[i] = static_cast<uint32_t>(l[2 * i]) | (static_cast<uint32_t>(l[2 * i + 1]) << 16);
    }
    *reinterpret_cast<u32x4*>(hi + g * 8) = ph;
    if (seg_lo >= 0) *reinterpret_cast<u32x4*>(lo + g * 8) = pl;
  }
}

// pack_blocks_kernel for a one-term pack (no remainder segment) of a column-major FP32 source
// with 16-byte aligned columns: each value is scaled and rounded as it is loaded, so the LDS
// transpose tile holds 16-bit values (17 KiB instead of 33): twice the workgroups share a CU,
// with twice the loads in flight.  Same block layout and the same values.
template <int FMT>
__global__ __launch_bounds__(256) void pack16_blocks_kernel(const float* __restrict__ src,
                                                            int64_t ld, int64_t rows, int64_t cols,
                                                            int64_t nkt, int nseg, int seg_hi,
                                                            const uint32_t* __restrict__ rowmax,
                                                            uint16_t* __restrict__ dst) {
  constexpr int LDT = kBR + 8;  // 16-bit values per LDS row (rows stay 16-byte aligned)
  __shared__ __attribute__((aligned(16))) uint16_t tile[kBKB][LDT];
  const int64_t tr = blockIdx.y, kt = blockIdx.x;
  const int64_t r0 = tr * kBR, k0 = kt * kBKB;
  const int t = threadIdx.x;
  const bool full = r0 + kBR <= rows;
  const int r = 4 * (t % 64);
  int e[4] = {0, 0, 0, 0};  // row scales (FP16)
  if (rowmax) {
#pragma unroll
    for (int i = 0; i < 4; ++i) e[i] = (r0 + r + i < rows) ? row_exponent(rowmax[r0 + r + i]) : 0;
  }
  float4 v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {  // a wave reads one 1 KiB column segment per instruction
    const int64_t gk = k0 + j * 4 + t / 64;
    v[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (gk < cols) {
      const float* col = src + gk * ld + r0 + r;
      if (full) {
        v[j] = *reinterpret_cast<const float4*>(col);
      } else {
        v[j].x = (r0 + r < rows) ? col[0] : 0.f;
        v[j].y = (r0 + r + 1 < rows) ? col[1] : 0.f;
        v[j].z = (r0 + r + 2 < rows) ? col[2] : 0.f;
        v[j].w = (r0 + r + 3 < rows) ? col[3] : 0.f;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = j * 4 + t / 64;
    const float x[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
    uint16_t h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = to16<FMT>(e[i] ? ldexpf(x[i], -e[i]) : x[i]);
    uint2 w;
    w.x = static_cast<uint32_t>(h[0]) | (static_cast<uint32_t>(h[1]) << 16);
    w.y = static_cast<uint32_t>(h[2]) | (static_cast<uint32_t>(h[3]) << 16);
    *reinterpret_cast<uint2*>(&tile[k][r]) = w;
  }
  __syncthreads();
  // every store instruction writes 1 KiB contiguous: chunk g = q*256 + t is (row g/4, slot g%4)
  uint16_t* hi = dst + ((tr * nseg + seg_hi) * nkt + kt) * kBlk;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int g = q * 256 + t;
    const int rr = g >> 2;
    const int c = (g & 3) ^ ((rr >> 2) & 3);
    u32x4 ph;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      ph[i] = static_cast<uint32_t>(tile[8 * c + 2 * i][rr]) |
              (static_cast<uint32_t>(tile[8 * c + 2 * i + 1][rr]) << 16);
    *reinterpret_cast<u32x4*>(hi + g * 8) = ph;
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

// Y (m x nl) = sum over segments s of A_seg(s) * B_seg(s), from blocks
// Ab[tm][a segment][kt] and Wb[tn][b segment][kt]: step kt of the K loop (0 <= kt < nseg * nkt)
// reads k block kt - s*nkt of A segment amap[s] and B segment bmap[s], s = kt / nkt (2 bits per
// segment in amap/bmap; ONESEG: a single segment, block kt).  256 x 256 tiles, 4 waves of
// 128 x 128 (v_mfma_f32_16x16x32_bf16/_f16, 256 accumulators per lane).  Blocks stream from HBM
// into LDS with global_load_lds_dwordx4 through a five-slot ring.  The ring buffers are distinct
// LDS arrays with static indices and the loop body has no branches (the stage four ahead is
// issued every step, clamped to the last block; the K step count is a multiple of 10), so the
// compiler keeps per-buffer wait tracking and the only waits are the explicit ones.
// rowmax (FP16): row i of Y is multiplied by 2^e_i.
// The grid covers row tiles [tm0, tm0 + tcount) (x: tcount x tiles_n tiles) and the K steps in
// ranges of kper (y: the range; kper a multiple of kUnroll).  P null: the sums go to Y; else to
// FP32 partials P[range][row - 256 tm0][column] (unscaled), summed by sketch_tail_reduce_kernel.
template <typename TY, int FMT, bool ONESEG>
__global__ __launch_bounds__(256) void sketch_blk_kernel(const uint16_t* __restrict__ Ab,
                                                        const uint16_t* __restrict__ Wb,
                                                        int64_t nkt, int nseg, int nseg_a,
                                                        int nseg_b, uint32_t amap, uint32_t bmap,
                                                        const uint32_t* __restrict__ rowmax,
                                                        int64_t m, int64_t nl,
                                                        TY* __restrict__ Y, int64_t ldy,
                                                        int64_t tm0, int64_t tcount, int64_t kper,
                                                        float* __restrict__ P) {
  constexpr int L = 8;  // DMA instructions per wave and stage (4 for A, 4 for B)
  static_assert(kStages == 5 && kUnroll == 10, "the loop body below is written for 5 slots");
  __shared__ __attribute__((aligned(16))) uint16_t S0[2 * kBlk], S1[2 * kBlk], S2[2 * kBlk],
      S3[2 * kBlk], S4[2 * kBlk];
  const int64_t tiles_n = ceil_div(nl, kBR);
  const int64_t logical = xcd_swizzle(blockIdx.x, tcount * tiles_n);
  const int64_t tm = tm0 + logical / tiles_n, tn = logical % tiles_n;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int wm0 = (wave / 2) * 128, wn0 = (wave % 2) * 128;
  const int kg = lane / 16;
  const int64_t kbeg = static_cast<int64_t>(blockIdx.y) * kper;
  const int64_t ktot = (kbeg + kper < nseg * nkt) ? kbeg + kper : nseg * nkt;  // this range's end
  const uint16_t* abase = Ab + tm * nseg_a * nkt * kBlk;
  const uint16_t* bbase = Wb + tn * nseg_b * nkt * kBlk;
  // the A and B blocks of step kt
  auto blocks = [&](int64_t kt, const uint16_t*& ak, const uint16_t*& bk) {
    if constexpr (ONESEG) {
      ak = abase + kt * kBlk;
      bk = bbase + kt * kBlk;
    } else {  // segment of step kt without a division: at most three segments
      const int s = static_cast<int>(kt >= nkt) + static_cast<int>(kt >= 2 * nkt);
      const int64_t kk = kt - s * nkt;
      ak = abase + (((amap >> (2 * s)) & 3) * nkt + kk) * kBlk;
      bk = bbase + (((bmap >> (2 * s)) & 3) * nkt + kk) * kBlk;
    }
  };
  // DMA instruction i of a stage (i / 2 = 1 KiB chunk of the wave, i % 2 = A or B)
  auto dma = [&](const uint16_t* ak, const uint16_t* bk, uint16_t* sd, int i) {
    const int off = ((i >> 1) * 4 + wave) * 512;  // 1 KiB per wave instruction
    if (i & 1)
      __builtin_amdgcn_global_load_lds(bk + off + lane * 8, (lds_void*)&sd[kBlk + off], 16, 0, 0);
    else
      __builtin_amdgcn_global_load_lds(ak + off + lane * 8, (lds_void*)&sd[off], 16, 0, 0);
  };
  auto issue = [&](int64_t kt, uint16_t* sd) {
    const uint16_t *ak, *bk;
    blocks(kt, ak, bk);
#pragma unroll
    for (int i = 0; i < L; ++i) dma(ak, bk, sd, i);
  };
  f4 acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // Fragments are double-buffered: step kt reads stage kt + 1's fragments from LDS while its
  // MFMAs consume stage kt's, which were read one step earlier.
  u32x4 fa0[8], fb0[8], fa1[8], fb1[8];
  auto frags = [&](const uint16_t* ss, u32x4* af, u32x4* bfr) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      af[i] = *reinterpret_cast<const u32x4*>(&ss[blk_off(wm0 + i * 16 + lane % 16, 8 * kg)]);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      bfr[j] = *reinterpret_cast<const u32x4*>(&ss[kBlk + blk_off(wn0 + j * 16 + lane % 16, 8 * kg)]);
  };
  // Step kt: stage kt + 1 must have landed (stages up to kt + 3 are issued; two may still be in
  // flight); the barrier makes that true in every wave.  The stage issued here, kt + 4 (clamped
  // to the last block), overwrites stage kt - 1's slot, whose fragments every wave consumed in
  // step kt - 1, before this barrier.  The last step's look-ahead reads a slot that holds the
  // last block again and discards it.  The step's MFMAs depend only on fragments already in
  // registers, so its memory instructions go between the rows of MFMAs (after row i: DMA
  // instruction i and stage kt + 1's fragments i of A and B), in an order fixed by compiler
  // fences: issued in one burst, 16 fragment reads and 8 DMA instructions stalled the wave's
  // MFMA issue behind the LDS queue (measured: the burst form took ~15 % longer).
  auto step = [&](int64_t kt, const u32x4* caf, const u32x4* cbf, const uint16_t* snext,
                  u32x4* naf, u32x4* nbf, uint16_t* sn) {
    asm volatile("" ::: "memory");
    wait_vmcnt_lgkm0<2 * L>();  // and this step's fragments (read last step) are in registers
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    const uint16_t *ak, *bk;
    blocks(kt + 4 < ktot ? kt + 4 : ktot - 1, ak, bk);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
#pragma unroll
      for (int j = 0; j < 8; ++j) mfma_tied<FMT>(acc[i][j], caf[i], cbf[j]);
      asm volatile("" ::: "memory");
      dma(ak, bk, sn, i);
      naf[i] = *reinterpret_cast<const u32x4*>(&snext[blk_off(wm0 + i * 16 + lane % 16, 8 * kg)]);
      nbf[i] = *reinterpret_cast<const u32x4*>(
          &snext[kBlk + blk_off(wn0 + i * 16 + lane % 16, 8 * kg)]);
      asm volatile("" ::: "memory");
    }
  };
  issue(kbeg, S0);
  issue(kbeg + 1, S1);
  issue(kbeg + 2, S2);
  issue(kbeg + 3, S3);
  asm volatile("" ::: "memory");
  wait_vmcnt<3 * L>();  // stage 0 landed in every wave
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  frags(S0, fa0, fb0);
  asm volatile("s_nop 7");  // zeroed accumulators (VALU writes) before their first MFMA read
  for (int64_t kt = kbeg; kt < ktot; kt += kUnroll) {  // stage s sits in slot s % 5 (kbeg % 10 == 0)
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
  if (P != nullptr) {
    float* Pr = P + (static_cast<int64_t>(blockIdx.y) * tcount * kBR - tm0 * kBR) * nl;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = m0 + wm0 + i * 16 + 4 * (lane / 16) + r;
        if (gi >= m) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int64_t gj = n0 + wn0 + j * 16 + lane % 16;
          if (gj < nl) Pr[gi * nl + gj] = acc[i][j][r];
        }
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t gi = m0 + wm0 + i * 16 + 4 * (lane / 16) + r;
      if (gi >= m) continue;
      const int e = rowmax ? row_exponent(rowmax[gi]) : 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t gj = n0 + wn0 + j * 16 + lane % 16;
        if (gj < nl) {
          if constexpr (sizeof(TY) == 8)
            Y[gi * ldy + gj] = ldexp(static_cast<double>(acc[i][j][r]), e);
          else
            Y[gi * ldy + gj] = ldexpf(acc[i][j][r], e);
        }
      }
    }
}

// Y rows [row0, row0 + rows) from the FP32 partials of sketch_blk_kernel's K ranges: their sum
// in range order (FP32, as the kernel accumulates), times 2^e_row for FP16.
template <typename TY>
__global__ __launch_bounds__(256) void sketch_tail_reduce_kernel(
    const float* __restrict__ P, int64_t ranges, int64_t prows, int64_t row0, int64_t rows,
    int64_t nl, const uint32_t* __restrict__ rowmax, TY* __restrict__ Y, int64_t ldy) {
  const int64_t total = rows * nl;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = e / nl, j = e % nl;
    float s = 0.f;
    for (int64_t q = 0; q < ranges; ++q) s += P[(q * prows + i) * nl + j];
    const int64_t gi = row0 + i;
    const int ex = rowmax ? row_exponent(rowmax[gi]) : 0;
    if constexpr (sizeof(TY) == 8)
      Y[gi * ldy + j] = ldexp(static_cast<double>(s), ex);
    else
      Y[gi * ldy + j] = ldexpf(s, ex);
  }
}

// Segment layout of a term count: A segments, B segments, and the (A, B) segment of each step.
struct Terms {
  int nseg, nseg_a, nseg_b;
  uint32_t amap, bmap;
};
Terms terms_of(int terms) {
  switch (terms) {
    case 2: return Terms{2, 2, 1, 0u | (1u << 2), 0u};                    // hi*W + lo*W
    case 3: return Terms{3, 2, 2, 0u | (1u << 2), 1u << 4};               // + hi*W_lo
    default: return Terms{1, 1, 1, 0u, 0u};
  }
}

struct BlkDims {
  int64_t tm, tn, nkt;
};
BlkDims blk_dims(const SketchBf16Args& a) {
  return BlkDims{ceil_div(a.m, kBR), ceil_div(a.nl, kBR), blk_nkt(a.n)};
}

template <typename TS, int FMT>
void launch_pack_fmt(const TS* src, int64_t ld, bool r_contig, int64_t rows, int64_t cols,
                     int64_t tiles, int64_t nkt, int nseg, int seg_hi, int seg_lo,
                     const uint32_t* rowmax, uint16_t* dst, hipStream_t st) {
  const dim3 grid(static_cast<unsigned>(nkt), static_cast<unsigned>(tiles));
  const bool vec4 = r_contig && ld % 4 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0;
  if constexpr (sizeof(TS) == 4) {
    if (vec4 && seg_lo < 0) {
      hipLaunchKernelGGL((pack16_blocks_kernel<FMT>), grid, dim3(256), 0, st, src, ld, rows, cols,
                         nkt, nseg, seg_hi, rowmax, dst);
      MCNLA_LAUNCHED();
      return;
    }
    if (vec4) {
      hipLaunchKernelGGL((pack_blocks_kernel<TS, true, true, FMT>), grid, dim3(256), 0, st, src,
                         ld, rows, cols, nkt, nseg, seg_hi, seg_lo, rowmax, dst);
      MCNLA_LAUNCHED();
      return;
    }
  }
  if (r_contig)
    hipLaunchKernelGGL((pack_blocks_kernel<TS, true, false, FMT>), grid, dim3(256), 0, st, src, ld,
                       rows, cols, nkt, nseg, seg_hi, seg_lo, rowmax, dst);
  else
    hipLaunchKernelGGL((pack_blocks_kernel<TS, false, false, FMT>), grid, dim3(256), 0, st, src,
                       ld, rows, cols, nkt, nseg, seg_hi, seg_lo, rowmax, dst);
  MCNLA_LAUNCHED();
}

template <typename TS>
void launch_pack(int fmt, const TS* src, int64_t ld, bool r_contig, int64_t rows, int64_t cols,
                 int64_t tiles, int64_t nkt, int nseg, int seg_hi, int seg_lo,
                 const uint32_t* rowmax, uint16_t* dst, hipStream_t st) {
  if (fmt == kFmtF16)
    launch_pack_fmt<TS, kFmtF16>(src, ld, r_contig, rows, cols, tiles, nkt, nseg, seg_hi, seg_lo,
                                 rowmax, dst, st);
  else
    launch_pack_fmt<TS, kFmtBf16>(src, ld, r_contig, rows, cols, tiles, nkt, nseg, seg_hi, seg_lo,
                                  rowmax, dst, st);
}

template <typename TS>
void launch_rowmax(const TS* A, int64_t lda, bool colmajor, int64_t m, int64_t n,
                   uint32_t* rowmax, hipStream_t st) {
  if (colmajor) {
    MCNLA_HIP(hipMemsetAsync(rowmax, 0, static_cast<size_t>(m) * sizeof(uint32_t), st));
    // K in chunks of at least 256, so that a CU holds enough rows in flight
    const int64_t rows_blocks = ceil_div(m, 256);
    const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(2048, rows_blocks),
                                                                  ceil_div(n, 256)));
    const int64_t kchunk = ceil_div(n, chunks);
    const dim3 grid(static_cast<unsigned>(rows_blocks), static_cast<unsigned>(ceil_div(n, kchunk)));
    hipLaunchKernelGGL(rowmax_colmajor_kernel<TS>, grid, dim3(256), 0, st, A, lda, m, n, kchunk,
                       rowmax);
  } else {
    hipLaunchKernelGGL(rowmax_rowmajor_kernel<TS>, dim3(static_cast<unsigned>(ceil_div(m, 4))),
                       dim3(256), 0, st, A, lda, m, n, rowmax);
  }
  MCNLA_LAUNCHED();
}

struct Workspace {
  uint16_t* Ab;
  uint16_t* Wb;
  uint32_t* rowmax;
  size_t bytes;
};
Workspace carve(const SketchBf16Args& a, void* base) {
  const BlkDims d = blk_dims(a);
  const Terms t = terms_of(a.terms);
  const size_t ab = static_cast<size_t>(t.nseg_a * d.tm * d.nkt) * kBlk * sizeof(uint16_t);
  const size_t wb = static_cast<size_t>(t.nseg_b * d.tn * d.nkt) * kBlk * sizeof(uint16_t);
  const size_t rb = a.fmt == kFmtF16 ? (static_cast<size_t>(a.m) * 4 + 255) / 256 * 256 : 0;
  char* p = static_cast<char*>(base);
  return Workspace{reinterpret_cast<uint16_t*>(p), reinterpret_cast<uint16_t*>(p + ab),
                   rb ? reinterpret_cast<uint32_t*>(p + ab + wb) : nullptr, ab + wb + rb};
}

// One workgroup per CU (160 KiB of LDS each): the row tiles that fill whole rounds of the grid
// run in full-length workgroups, the rest (at most half a round) with K split into `ranges`
// ranges of kper steps, so that their last round ends sooner.
struct TailPlan {
  int64_t main_tm, tail_tm, ranges, kper;
};
TailPlan tail_plan(const BlkDims& d, const Terms& t) {
  const int64_t ktot = t.nseg * d.nkt;
  TailPlan p{d.tm, 0, 1, ktot};
  const int64_t cus = compute_units();
  int64_t g = cus, b = d.tn;
  while (b != 0) {
    const int64_t r = g % b;
    g = b;
    b = r;
  }
  const int64_t per_round = cus / g;  // row tiles whose tiles fill whole rounds
  const int64_t main_tm = d.tm / per_round * per_round, tail = (d.tm - main_tm) * d.tn;
  if (main_tm == 0 || tail == 0 || 2 * tail > cus) return p;
  const int64_t units = ktot / kUnroll;  // K step groups
  int64_t ranges = cus / tail;
  if (ranges > 4) ranges = 4;
  if (ranges > units / 4) ranges = units / 4;  // each range at least 4 groups
  if (ranges < 2) return p;
  const int64_t kper = ceil_div(units, ranges) * kUnroll;
  return TailPlan{main_tm, d.tm - main_tm, ceil_div(ktot, kper), kper};
}

}  // namespace

size_t sketch_bf16_workspace(const SketchBf16Args& a) { return carve(a, nullptr).bytes; }

size_t sketch_bf16_tail_bytes(const SketchBf16Args& a) {
  if (a.m == 0 || a.nl == 0) return 0;
  const TailPlan p = tail_plan(blk_dims(a), terms_of(a.terms));
  return p.tail_tm ? static_cast<size_t>(p.ranges * p.tail_tm * kBR * a.nl) * sizeof(float) : 0;
}

void sketch_bf16(const SketchBf16Args& a, void* workspace, size_t workspace_bytes, hipStream_t st,
                 hipEvent_t gemm_begin, hipEvent_t gemm_end) {
  MCNLA_REQUIRE(a.terms >= 1 && a.terms <= 3, MCNLA_ERR_INVALID, "terms must be 1, 2 or 3");
  MCNLA_REQUIRE(a.fmt == kFmtBf16 || a.fmt == kFmtF16, MCNLA_ERR_INVALID, "unknown 16-bit format");
  MCNLA_REQUIRE(workspace_bytes >= sketch_bf16_workspace(a), MCNLA_ERR_INVALID,
                "16-bit sketch workspace too small");
  if (a.m == 0 || a.nl == 0) return;
  const BlkDims d = blk_dims(a);
  const Terms t = terms_of(a.terms);
  const Workspace w = carve(a, workspace);
  // A: element (i, k) at A[i + k*lda] (col-major) or A[i*lda + k] (row-major); hi in segment 0,
  // lo in segment 1 for two or three terms.
  const int a_lo = t.nseg_a == 2 ? 1 : -1;
  if (a.a_f32) {
    const float* A = static_cast<const float*>(a.A);
    if (w.rowmax) launch_rowmax(A, a.lda, a.a_colmajor, a.m, a.n, w.rowmax, st);
    launch_pack(a.fmt, A, a.lda, a.a_colmajor, a.m, a.n, d.tm, d.nkt, t.nseg_a, 0, a_lo, w.rowmax,
                w.Ab, st);
  } else {
    const double* A = static_cast<const double*>(a.A);
    if (w.rowmax) launch_rowmax(A, a.lda, a.a_colmajor, a.m, a.n, w.rowmax, st);
    launch_pack(a.fmt, A, a.lda, a.a_colmajor, a.m, a.n, d.tm, d.nkt, t.nseg_a, 0, a_lo, w.rowmax,
                w.Ab, st);
  }
  // Omega^T: element (j, k) = Omega[k*nl + j], i.e. row index j contiguous; N(0,1) values need
  // no scaling in either format.
  launch_pack(a.fmt, a.omega, a.nl, true, a.nl, a.n, d.tn, d.nkt, t.nseg_b, 0,
              t.nseg_b == 2 ? 1 : -1, static_cast<const uint32_t*>(nullptr), w.Wb, st);
  TailPlan tp{d.tm, 0, 1, t.nseg * d.nkt};
  if (a.tail != nullptr) {
    const TailPlan q = tail_plan(d, t);
    if (q.tail_tm && a.tail_bytes >= static_cast<size_t>(q.ranges * q.tail_tm * kBR * a.nl) *
                                          sizeof(float))
      tp = q;
  }
  if (gemm_begin) MCNLA_HIP(hipEventRecord(gemm_begin, st));
  auto launch = [&](auto* Y, auto fmt_tag) {
    using TY = std::remove_pointer_t<decltype(Y)>;
    constexpr int FMT = decltype(fmt_tag)::value;
    const uint32_t* rm = w.rowmax;
    // tm0, tcount, kper, partials: the rows of whole rounds, then the rest in K ranges
    auto gemm = [&](int64_t tm0, int64_t tcount, int64_t ranges, int64_t kper, float* P) {
      const dim3 grid(static_cast<unsigned>(tcount * d.tn), static_cast<unsigned>(ranges));
      if (t.nseg == 1)
        hipLaunchKernelGGL((sketch_blk_kernel<TY, FMT, true>), grid, dim3(256), 0, st, w.Ab, w.Wb,
                           d.nkt, t.nseg, t.nseg_a, t.nseg_b, t.amap, t.bmap, rm, a.m, a.nl, Y,
                           a.ldy, tm0, tcount, kper, P);
      else
        hipLaunchKernelGGL((sketch_blk_kernel<TY, FMT, false>), grid, dim3(256), 0, st, w.Ab,
                           w.Wb, d.nkt, t.nseg, t.nseg_a, t.nseg_b, t.amap, t.bmap, rm, a.m, a.nl,
                           Y, a.ldy, tm0, tcount, kper, P);
      MCNLA_LAUNCHED();
    };
    gemm(0, tp.main_tm, 1, t.nseg * d.nkt, nullptr);
    if (tp.tail_tm) {
      gemm(tp.main_tm, tp.tail_tm, tp.ranges, tp.kper, a.tail);
      const int64_t row0 = tp.main_tm * kBR, rows = a.m - row0;
      const int64_t blocks = std::min<int64_t>(ceil_div(rows * a.nl, 256), 4096);
      hipLaunchKernelGGL(sketch_tail_reduce_kernel<TY>, dim3(static_cast<unsigned>(blocks)),
                         dim3(256), 0, st, a.tail, tp.ranges, tp.tail_tm * kBR, row0, rows, a.nl,
                         rm, Y, a.ldy);
    }
  };
  using Bf = std::integral_constant<int, kFmtBf16>;
  using F16 = std::integral_constant<int, kFmtF16>;
  if (a.Yd) {
    if (a.fmt == kFmtF16) launch(a.Yd, F16{}); else launch(a.Yd, Bf{});
  } else {
    if (a.fmt == kFmtF16) launch(a.Y, F16{}); else launch(a.Y, Bf{});
  }
  MCNLA_LAUNCHED();
  if (gemm_end) MCNLA_HIP(hipEventRecord(gemm_end, st));
}

}  // namespace mcnla
