// FP64 MFMA GEMM for gfx950: C = alpha * op(A) * op(B) + beta * C, with split-K and an optional
// lower-triangle tile schedule for Gram matrices.
//
// Operands are described by which index is contiguous in memory, so one kernel covers every
// dense contraction of the iSVD path:
//   sketch        Y  = A * Omega        A col-major (M-contiguous) or row-major (K-contiguous),
//                                       Omega row-major (N-contiguous)
//   Gram          G  = Y' * Y           both operands are rows of Y: M- and N-contiguous
//   basis         Qb = Y * M            Y row-major (K-contiguous), M row-major
//   former        Z  = A' * Qbar        A col-major (K-contiguous) / row-major (M-contiguous)
// Tiles are staged through LDS as [k][mn] (mn contiguous) so that every MFMA operand fragment is
// one LDS row segment; v_mfma_f64_16x16x4f64 does the arithmetic.  Source elements may be FP32
// (mixed precision collections); they are widened to FP64 on the way into LDS.
#include <atomic>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"
#include "tall_direct.hpp"

namespace mcnla {

namespace {

template <int BM, int BN, int BK, int WM, int WN>
struct TileCfg {
  static constexpr int kWavesM = BM / WM;
  static constexpr int kWavesN = BN / WN;
  static constexpr int kThreads = kWavesM * kWavesN * kWave;
  static constexpr int kFragM = WM / 16;
  static constexpr int kFragN = WN / 16;
  static constexpr int kPadA = 4;  // doubles of padding per LDS row
  static constexpr int kPadB = 4;
  static constexpr int kLdA = BM + kPadA;
  static constexpr int kLdB = BN + kPadB;
  static constexpr int kElemsA = BM * BK / kThreads;
  static constexpr int kElemsB = BK * BN / kThreads;
  static_assert(BM % WM == 0 && BN % WN == 0 && WM % 16 == 0 && WN % 16 == 0, "tile shape");
  static_assert((BM * BK) % kThreads == 0 && (BK * BN) % kThreads == 0, "load split");
  static_assert(BK % 4 == 0, "BK");
};

// Thread -> element map of an operand tile (BK k-values x MN values).  Elements go in groups of
// VEC that are consecutive along the contiguous index: group g = e * NT + tid of a thread holds
// (k, mn..mn+VEC-1) when MN_CONTIG, else (k..k+VEC-1, mn).  Loads and LDS stores use the same map.
template <bool MN_CONTIG, int MN, int BK, int VEC>
__device__ __forceinline__ void group_origin(int g, int& kk, int& mm) {
  if (MN_CONTIG) {
    kk = g / (MN / VEC);
    mm = (g % (MN / VEC)) * VEC;
  } else {
    mm = g / (BK / VEC);
    kk = (g % (BK / VEC)) * VEC;
  }
}

// Loads one operand tile into registers.  MN_CONTIG: element (k, mn) lives at base[mn + k*ld];
// otherwise at base[mn*ld + k].  A full tile with VEC > 1 is read with one 8/16-byte load per
// group (the launcher guarantees the alignment); partial tiles are read element by element with
// zero fill.
template <typename TS, bool MN_CONTIG, int MN, int BK, int NT, int ELEMS, int VEC>
__device__ __forceinline__ void load_tile(double (&r)[ELEMS], const TS* __restrict__ base,
                                          int64_t ld, int64_t mn0, int64_t k0, int64_t mn_lim,
                                          int64_t k_lim, bool full) {
  static_assert(ELEMS % VEC == 0, "vector groups");
  const int tid = threadIdx.x;
#pragma unroll
  for (int e = 0; e < ELEMS / VEC; ++e) {
    int kk, mm;
    group_origin<MN_CONTIG, MN, BK, VEC>(e * NT + tid, kk, mm);
    const int64_t gm = mn0 + mm, gk = k0 + kk;
    bool done = false;
    if constexpr (VEC > 1) {
      if (full) {
        typedef typename VecOf<TS, VEC>::type V;
        const int64_t off = MN_CONTIG ? (gm + gk * ld) : (gm * ld + gk);
        const V v = *reinterpret_cast<const V*>(base + off);
#pragma unroll
        for (int q = 0; q < VEC; ++q) r[e * VEC + q] = static_cast<double>(v[q]);
        done = true;
      }
    }
    if (!done) {
#pragma unroll
      for (int q = 0; q < VEC; ++q) {
        const int64_t qm = MN_CONTIG ? gm + q : gm, qk = MN_CONTIG ? gk : gk + q;
        r[e * VEC + q] = (full || (qm < mn_lim && qk < k_lim))
                             ? static_cast<double>(base[MN_CONTIG ? (qm + qk * ld) : (qm * ld + qk)])
                             : 0.0;
      }
    }
  }
}

template <bool MN_CONTIG, int MN, int BK, int NT, int ELEMS, int LD, int VEC>
__device__ __forceinline__ void store_tile(double* __restrict__ lds, const double (&r)[ELEMS]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int e = 0; e < ELEMS / VEC; ++e) {
    int kk, mm;
    group_origin<MN_CONTIG, MN, BK, VEC>(e * NT + tid, kk, mm);
#pragma unroll
    for (int q = 0; q < VEC; ++q) {
      if (MN_CONTIG)
        lds[kk * LD + mm + q] = r[e * VEC + q];
      else
        lds[(kk + q) * LD + mm] = r[e * VEC + q];
    }
  }
}

constexpr int vec_for(int elem_bytes, int elems, bool ok) {
  return !ok ? 1 : ((16 / elem_bytes) <= elems ? 16 / elem_bytes : elems);
}

template <typename TA, typename TB, bool A_MC, bool B_NC, int BM, int BN, int BK, int WM, int WN,
          bool VECOK>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_f64_kernel(
    GemmF64Args p) {
  using Cfg = TileCfg<BM, BN, BK, WM, WN>;
  constexpr int NT = Cfg::kThreads;
  constexpr int VA = vec_for(sizeof(TA), Cfg::kElemsA, VECOK);
  constexpr int VB = vec_for(sizeof(TB), Cfg::kElemsB, VECOK);
  __shared__ double As[2][BK * Cfg::kLdA];
  __shared__ double Bs[2][BK * Cfg::kLdB];

  if (p.gate != nullptr && *p.gate < p.gate_t) return;
  // ---- tile schedule: logical id = (batch, split, tile), tile fastest -------------------------
  const int64_t tiles_m = ceil_div(p.M, BM), tiles_n = ceil_div(p.N, BN);
  const int64_t ntiles = p.lower_only ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
  const int64_t nblocks = ntiles * p.splits * p.batch;
  const int64_t logical = xcd_swizzle(blockIdx.x, nblocks);
  const int64_t bat = logical / (ntiles * p.splits);
  const int64_t split = (logical / ntiles) % p.splits;
  const int64_t t = logical % ntiles;

  const TA* __restrict__ A = static_cast<const TA*>(p.A) + bat * p.stride_a;
  const TB* __restrict__ B = static_cast<const TB*>(p.B) + bat * p.stride_b;
  int64_t tm, tn;
  if (p.lower_only) {
    tm = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(t) + 1.0) - 1.0) * 0.5);
    while (tm * (tm + 1) / 2 > t) --tm;
    while ((tm + 1) * (tm + 2) / 2 <= t) ++tm;
    tn = t - tm * (tm + 1) / 2;
  } else {
    tm = t / tiles_n;
    tn = t % tiles_n;
  }
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;

  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int wm0 = (wave / Cfg::kWavesN) * WM;
  const int wn0 = (wave % Cfg::kWavesN) * WN;

  d4 acc[Cfg::kFragM][Cfg::kFragN];
#pragma unroll
  for (int i = 0; i < Cfg::kFragM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::kFragN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  const bool full_mn = (m0 + BM <= p.M) && (n0 + BN <= p.N);
  double ra[Cfg::kElemsA], rb[Cfg::kElemsB];

  if (kbeg < kend) {
    const int64_t ntk = ceil_div(kend - kbeg, BK);
    {
      const bool full = full_mn && (kbeg + BK <= kend);
      load_tile<TA, A_MC, BM, BK, NT, Cfg::kElemsA, VA>(ra, A, p.lda, m0, kbeg, p.M, kend, full);
      load_tile<TB, B_NC, BN, BK, NT, Cfg::kElemsB, VB>(rb, B, p.ldb, n0, kbeg, p.N, kend, full);
      store_tile<A_MC, BM, BK, NT, Cfg::kElemsA, Cfg::kLdA, VA>(As[0], ra);
      store_tile<B_NC, BN, BK, NT, Cfg::kElemsB, Cfg::kLdB, VB>(Bs[0], rb);
    }
    __syncthreads();
    for (int64_t kt = 0; kt < ntk; ++kt) {
      const int cur = kt & 1;
      const bool more = kt + 1 < ntk;
      if (more) {
        const int64_t k0 = kbeg + (kt + 1) * BK;
        const bool full = full_mn && (k0 + BK <= kend);
        load_tile<TA, A_MC, BM, BK, NT, Cfg::kElemsA, VA>(ra, A, p.lda, m0, k0, p.M, kend, full);
        load_tile<TB, B_NC, BN, BK, NT, Cfg::kElemsB, VB>(rb, B, p.ldb, n0, k0, p.N, kend, full);
      }
      const double* __restrict__ as = As[cur];
      const double* __restrict__ bs = Bs[cur];
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int krow = kk + lane / 16;
        double af[Cfg::kFragM], bf[Cfg::kFragN];
#pragma unroll
        for (int i = 0; i < Cfg::kFragM; ++i) af[i] = as[krow * Cfg::kLdA + wm0 + i * 16 + lane % 16];
#pragma unroll
        for (int j = 0; j < Cfg::kFragN; ++j) bf[j] = bs[krow * Cfg::kLdB + wn0 + j * 16 + lane % 16];
#pragma unroll
        for (int i = 0; i < Cfg::kFragM; ++i)
#pragma unroll
          for (int j = 0; j < Cfg::kFragN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
      if (more) {
        store_tile<A_MC, BM, BK, NT, Cfg::kElemsA, Cfg::kLdA, VA>(As[cur ^ 1], ra);
        store_tile<B_NC, BN, BK, NT, Cfg::kElemsB, Cfg::kLdB, VB>(Bs[cur ^ 1], rb);
      }
      __syncthreads();
    }
  }

  // ---- epilogue: for v_mfma_f64_16x16x4f64 lane holds D[lane/16 + 4*r][lane%16] of each
  // 16x16 fragment (rows interleaved by 4, unlike the f32 16x16 MFMAs) -----------------------
  double* __restrict__ C = p.partial ? p.partial + (bat * p.splits + split) * p.M * p.N
                                     : p.C + bat * p.stride_c;
  const int64_t ldc = p.partial ? p.N : p.ldc;
#pragma unroll
  for (int i = 0; i < Cfg::kFragM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::kFragN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = m0 + wm0 + i * 16 + (lane / 16) + 4 * r;
        const int64_t gj = n0 + wn0 + j * 16 + lane % 16;
        if (gi < p.M && gj < p.N) {
          const double v = acc[i][j][r];
          if (p.partial) {
            C[gi * ldc + gj] = v;
          } else {
            double* c = C + gi * ldc + gj;
            *c = (p.beta == 0.0) ? p.alpha * v : p.alpha * v + p.beta * *c;
          }
        }
      }
}

// ------------------------------------------------------------------------------------------------
// Fast path for FP64 operands that are both contiguous along M / N (the sketch of a column-major
// A, and Gram matrices Y'Y of a row-major Y): tiles go from HBM straight into LDS with
// global_load_lds_dwordx4 (no staging registers), through a three-stage ring.
//
// Each wave instruction moves one 1 KiB row segment (128 doubles) of a tile.  Rows beyond M / N
// read from the tile's first row (always valid) and only feed output rows that are never stored;
// k-rows beyond the split's range read its first k-row and are zeroed in the A fragments, so a
// tile never reads outside the operands.  Stage buffers are distinct LDS arrays addressed with
// static indices (the loop is unrolled by three), which lets the compiler see that the in-flight
// DMA of one stage never aliases the fragment reads of another; completion is tracked by hand
// with s_waitcnt vmcnt and a bare s_barrier.
typedef __attribute__((address_space(3))) void lds_void;

// One 16-byte-per-lane HBM -> LDS transfer, as a plain (non-template) function: the builtin
// called with template-dependent operand types makes the host pass drop the kernel's launch stub.
__device__ __forceinline__ void dma16(const void* g, lds_void* l) {
  __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
}

template <int VM>
__device__ __forceinline__ void wait_vmcnt() {
  // gfx9 s_waitcnt fields: vmcnt[3:0], expcnt[6:4], lgkmcnt[11:8], vmcnt high bits [15:14]
  __builtin_amdgcn_s_waitcnt((VM & 0xF) | (0x7 << 4) | (0xF << 8) | (((VM >> 4) & 3) << 14));
}

// A 64-bit value that is the same in every lane, moved into scalar registers.
__device__ __forceinline__ int64_t uniform64(int64_t x) {
  const uint64_t u = static_cast<uint64_t>(x);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

template <int BM, int BN, int BK, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_f64_dma_kernel(GemmF64Args p) {
  constexpr int NW = (BM / WM) * (BN / WN);
  // LDS row strides of 128 bytes modulo 256: a fragment read (lanes 0-15 on k-row r, 16-31 on
  // r + 1, 8 bytes each) then covers the 64 banks once per 32 lanes.  NARROW (BN = 64): B rows
  // are 512 bytes, unpadded, two k-rows per DMA instruction, and element (k, j) sits at
  // k BN + (j ^ 16 (k & 1)): odd rows have their 128-byte halves swapped, to the same effect.
  constexpr bool NARROW = BN == 64;
  constexpr int LDA = BM + 16, LDB = NARROW ? BN : BN + 16;
  constexpr int RA = BM / 128, RB = NARROW ? 1 : BN / 128;  // wave instructions per tile row
  constexpr int LA = BK * RA / NW, LB = (NARROW ? BK / 2 : BK * RB) / NW;  // per wave and stage
  constexpr int L = LA + LB;
  constexpr int FM = WM / 16, FN = WN / 16;
  static_assert(BM % 128 == 0 && (NARROW || BN % 128 == 0) && (BK * RA) % NW == 0 &&
                    (NARROW ? (BK / 2) % NW == 0 : (BK * RB) % NW == 0) && LB >= 1,
                "DMA tile shape");
  __shared__ __attribute__((aligned(16))) double A0[BK * LDA];
  __shared__ __attribute__((aligned(16))) double A1[BK * LDA];
  __shared__ __attribute__((aligned(16))) double A2[BK * LDA];
  __shared__ __attribute__((aligned(16))) double B0[BK * LDB];
  __shared__ __attribute__((aligned(16))) double B1[BK * LDB];
  __shared__ __attribute__((aligned(16))) double B2[BK * LDB];

  if (p.gate != nullptr && *p.gate < p.gate_t) return;
  constexpr int64_t R = BM / BN;  // the lower-triangle schedule needs R = 2 (host-checked)
  const int64_t tiles_m = ceil_div(p.M, BM), tiles_n = ceil_div(p.N, BN);
  int64_t bat, split, tm, tn;
  if (p.lower_only) {
    // Tile row tm of a symmetric output holds the 2 tm + 1 tiles left of the diagonal 256-block's
    // right half (full work, tm^2 before row tm) and that right half, whose upper waves skip their
    // MFMAs (about half the work).  The full tiles of every split come first and the half tiles
    // last, each group spread over the XCDs, so the short workgroups fill the end of the grid.
    const int64_t nf = tiles_m * tiles_m, nh = tiles_m;
    const int64_t fb = nf * p.splits * p.batch, hb = nh * p.splits * p.batch;
    const int64_t b = blockIdx.x;
    if (b < fb) {
      const int64_t lg = xcd_swizzle(b, fb);
      bat = lg / (nf * p.splits);
      split = (lg / nf) % p.splits;
      const int64_t t = lg % nf;
      tm = static_cast<int64_t>(sqrt(static_cast<double>(t)));
      while (tm * tm > t) --tm;
      while ((tm + 1) * (tm + 1) <= t) ++tm;
      tn = t - tm * tm;
    } else {
      const int64_t lg = xcd_swizzle(b - fb, hb);
      bat = lg / (nh * p.splits);
      split = (lg / nh) % p.splits;
      tm = lg % nh;
      tn = R * tm + 1;
      if (tn * BN >= p.N) return;  // the diagonal block's right half lies beyond the last column
    }
  } else {
    const int64_t ntiles = tiles_m * tiles_n;
    const int64_t lg = xcd_swizzle(blockIdx.x, ntiles * p.splits * p.batch);
    bat = lg / (ntiles * p.splits);
    split = (lg / ntiles) % p.splits;
    const int64_t t = lg % ntiles;
    tm = t / tiles_n;
    tn = t % tiles_n;
  }
  // The tile's coordinates are uniform by construction; moved into scalar registers (the
  // schedule's square root and 64-bit divisions leave them in vector registers), every DMA
  // address below is scalar arithmetic.
  const int64_t m0 = uniform64(tm * BM), n0 = uniform64(tn * BN);
  const int64_t kbeg = uniform64(split * p.kchunk);
  const int64_t kend = uniform64((kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K);
  const int lane = threadIdx.x % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kWave));
  const int wm0 = (wave / (BN / WN)) * WM, wn0 = (wave % (BN / WN)) * WN;
  const double* __restrict__ A = static_cast<const double*>(p.A) + uniform64(bat * p.stride_a);
  const double* __restrict__ B = static_cast<const double*>(p.B) + uniform64(bat * p.stride_b);
  // A wave whose WM x WN block lies strictly above the diagonal of a symmetric output only moves
  // tiles for the others; the reduction never reads its entries.
  const bool live = !p.lower_only || m0 + wm0 + WM - 1 >= n0 + wn0;

  d4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  // k-rows within the split are counted from kbeg in 32 bits (SALU has no signed 64-bit compare,
  // so 64-bit bounds checks would run on the VALU)
  const int krem = static_cast<int>(kend - kbeg);  // k-rows of this split
  const int ntk = krem > 0 ? (krem + BK - 1) / BK : 0;
  // DMA instruction q of a stage moves 128 doubles of k-row kk(q) of a tile, from column chunk
  // ch(q).  Its source is a uniform k-row address (the stage's first k-row plus kk(q) rows, or the
  // split's first k-row past the split's end) plus the lane's byte offset in the chunk, which is
  // fixed (columns past M / N read the tile's first column).
  const char* a0 = reinterpret_cast<const char*>(A + m0);
  const char* b0 = reinterpret_cast<const char*>(B + n0);
  const int64_t lda8 = p.lda * 8, ldb8 = p.ldb * 8;
  uint32_t aoff[LA], boff[LB];
  int64_t akoff[LA], bkoff[LB];
#pragma unroll
  for (int q = 0; q < LA; ++q) {
    const int idx = q * NW + wave;
    const int64_t gm = m0 + (idx % RA) * 128 + lane * 2;
    aoff[q] = static_cast<uint32_t>(((gm < p.M) ? gm - m0 : 0) * 8);
    akoff[q] = (idx / RA) * lda8;
  }
#pragma unroll
  for (int q = 0; q < LB; ++q) {
    const int idx = q * NW + wave;
    if constexpr (NARROW) {  // k-rows 2 idx (lanes 0-31) and 2 idx + 1 (lanes 32-63)
      const int h = lane / 32;
      const int64_t gn = n0 + ((2 * (lane % 32)) ^ (16 * h));
      boff[q] = static_cast<uint32_t>(h * ldb8 + ((gn < p.N) ? gn - n0 : 0) * 8);
      bkoff[q] = 2 * idx * ldb8;
    } else {
      const int64_t gn = n0 + (idx % RB) * 128 + lane * 2;
      boff[q] = static_cast<uint32_t>(((gn < p.N) ? gn - n0 : 0) * 8);
      bkoff[q] = (idx / RB) * ldb8;
    }
  }
  // DMA instruction i (A for i < LA, then B) of the stage starting at k-row k0
  const char* a_first = a0 + kbeg * lda8;  // the split's first k-row
  const char* b_first = b0 + kbeg * ldb8;
  // DMA instruction i (A for i < LA, then B) of the stage starting at k-row k0, whose k-row
  // addresses start at arow / brow
  auto dma = [&](int i, int kofs, const char* arow, const char* brow, double* Ad, double* Bd) {
    if (i < LA) {
      const int q = i, kk = (q * NW + wave) / RA, ch = (q * NW + wave) % RA;
      const char* row = (kofs + kk < krem) ? arow + akoff[q] : a_first;
      dma16(row + aoff[q], (lds_void*)&Ad[kk * LDA + ch * 128]);
    } else if constexpr (NARROW) {  // K even: a pair of k-rows is in range or out together
      const int q = i - LA, kk = 2 * (q * NW + wave);
      const char* row = (kofs + kk < krem) ? brow + bkoff[q] : b_first;
      dma16(row + boff[q], (lds_void*)&Bd[kk * LDB]);
    } else {
      const int q = i - LA, kk = (q * NW + wave) / RB, ch = (q * NW + wave) % RB;
      const char* row = (kofs + kk < krem) ? brow + bkoff[q] : b_first;
      dma16(row + boff[q], (lds_void*)&Bd[kk * LDB + ch * 128]);
    }
  };
  auto issue = [&](int kt, double* Ad, double* Bd) {
    const int kofs = kt * BK;
    const char* arow = a_first + kofs * lda8;
    const char* brow = b_first + kofs * ldb8;
#pragma unroll
    for (int i = 0; i < L; ++i) dma(i, kofs, arow, brow, Ad, Bd);
  };
  // Branch-free pipeline: the stage two ahead is issued every step (clamped to the last stage,
  // whose rows are then read again and never used), and the step count is padded to a multiple
  // of three with steps whose A fragments are all zero.  Without branches in the loop the
  // compiler keeps its per-buffer wait tracking, so the only DMA waits are the explicit ones.
  // MASK: the step may hold k-rows past the split's range (its last stage, and the padding
  // steps); the steps before it run without the masks.
  auto step = [&](auto mask, int kt, const double* as, const double* bs, double* An, double* Bn) {
    constexpr bool MASK = decltype(mask)::value;
    asm volatile("" ::: "memory");
    wait_vmcnt<L>();               // this stage landed; the next may still be in flight
    __builtin_amdgcn_s_barrier();  // ... for every wave; and stage kt-1 is fully read
    asm volatile("" ::: "memory");
    const int kn = kt + 2 < ntk ? kt + 2 : ntk - 1;
    const int left = krem - kt * BK;
    const int kvalid = left < BK ? left : BK;  // <= 0 past the last stage
    // The stage two ahead is issued DPG instructions at a time between the MFMAs of the first
    // k groups (its buffer was last read in step kt - 1, before the barrier): issued as one burst
    // after the barrier, the DMA held every wave's first MFMAs back.
    constexpr int NG = BK / 4, DPG = (L + NG - 2) / (NG - 1);
    const int kofs_n = kn * BK;
    const char* arow_n = a_first + kofs_n * lda8;
    const char* brow_n = b_first + kofs_n * ldb8;
    // A fragments of k-rows past the split's range are zeroed where an MFMA takes them, not as
    // they are read: a select on the read result would wait for the LDS read at once.
    double af[2][FM], bf[2][FN];
    bool aok[2];
    auto frags = [&](int kk, int b) {
      const int krow = kk + lane / 16;
#pragma unroll
      for (int i = 0; i < FM; ++i) af[b][i] = as[krow * LDA + wm0 + i * 16 + lane % 16];
      aok[b] = krow < kvalid;
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bf[b][j] = bs[krow * LDB + ((wn0 + j * 16 + lane % 16) ^ (NARROW ? 16 * (krow & 1) : 0))];
    };
    frags(0, 0);
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int b = (kk / 4) & 1;
      if (kk + 4 < BK) frags(kk + 4, b ^ 1);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              (!MASK || aok[b]) ? af[b][i] : 0.0, bf[b][j], acc[i][j], 0, 0, 0);
      if (kk + 4 < BK) {
        const int g = kk / 4;
#pragma unroll
        for (int i = g * DPG; i < (g + 1) * DPG && i < L; ++i) dma(i, kofs_n, arow_n, brow_n, An, Bn);
        // the next group's fragment reads (paired into (FM + FN) / 2 instructions) and this
        // group's DMA instructions, spread evenly between its FM FN MFMAs, MFMAs first: LDS DMA
        // shares the LDS counter, so each wait for fragments waits for every LDS read issued
        // before it, and a read issued ahead of the group's first MFMA would be waited for there
        constexpr int NQ = (FM + FN) / 2;
        const int nd = (g + 1) * DPG <= L ? DPG : (L > g * DPG ? L - g * DPG : 0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, FM * FN / NQ / 2, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
          if (q < nd) __builtin_amdgcn_sched_group_barrier(0x010, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, FM * FN / NQ - FM * FN / NQ / 2, 0);
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
  };
  // A wave that computes nothing runs the same barriers and DMA in a loop of its own (a branch
  // inside the loop would merge two DMA schedules, after which the compiler waits for every
  // DMA in flight before the next fragment read).
  auto move = [&](int kt, double* An, double* Bn) {
    asm volatile("" ::: "memory");
    wait_vmcnt<L>();
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    issue(kt + 2 < ntk ? kt + 2 : ntk - 1, An, Bn);
  };
  if (ntk > 0) {
    issue(0, A0, B0);
    issue(ntk > 1 ? 1 : 0, A1, B1);
    const int nsteps = (ntk + 2) / 3 * 3;
    if (live) {
      // whole stages of the split's range, three at a time, without masks; then the rest
      const int nfull = (krem / BK) / 3 * 3;
      const std::integral_constant<bool, false> unmasked;
      const std::integral_constant<bool, true> masked;
      int kt = 0;
      for (; kt < nfull; kt += 3) {
        step(unmasked, kt, A0, B0, A2, B2);
        step(unmasked, kt + 1, A1, B1, A0, B0);
        step(unmasked, kt + 2, A2, B2, A1, B1);
      }
      for (; kt < nsteps; kt += 3) {
        step(masked, kt, A0, B0, A2, B2);
        step(masked, kt + 1, A1, B1, A0, B0);
        step(masked, kt + 2, A2, B2, A1, B1);
      }
    } else {
      for (int kt = 0; kt < nsteps; kt += 3) {
        move(kt, A2, B2);
        move(kt + 1, A0, B0);
        move(kt + 2, A1, B1);
      }
    }
    wait_vmcnt<0>();  // no DMA may still target LDS when the workgroup ends
  }

  double* __restrict__ C = p.partial ? p.partial + (bat * p.splits + split) * p.M * p.N
                                     : p.C + bat * p.stride_c;
  const int64_t ldc = p.partial ? p.N : p.ldc;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = m0 + wm0 + i * 16 + (lane / 16) + 4 * r;
        const int64_t gj = n0 + wn0 + j * 16 + lane % 16;
        if (gi < p.M && gj < p.N) {
          const double v = acc[i][j][r];
          if (p.partial) {
            C[gi * ldc + gj] = v;
          } else {
            double* c = C + gi * ldc + gj;
            *c = (p.beta == 0.0) ? p.alpha * v : p.alpha * v + p.beta * *c;
          }
        }
      }
}

// ------------------------------------------------------------------------------------------------
// Tall-skinny products C (M x N, N <= 64) = A * B where A's rows run along K (element (i, k) at
// A[i*lda + k]) and B's rows along N (element (k, j) at B[k*ldb + j]): the former's Z = A'Qbar
// of a column-major A, Qbar = A X of a row-major A.  A is streamed once, so the kernel is bound
// by HBM or by MFMA, whichever is tighter.  This LDS-staged form serves FP32 A operands; FP64
// ones (Qbar = Y (T M), the integrators' Bs products) go to tall_direct_kernel below.
//
// 128 x LP output tiles (LP = 32 or 64), four waves of 32 x LP.  Each step moves a 128 x 32 slab
// of A (two threads per row, 16 consecutive k each: 128 or 256 contiguous bytes per row) and a
// 32 x LP slab of B through LDS in FP64, k-major, with row strides of 128 bytes modulo 256 so
// that the fragment reads of 32 lanes cover the 64 banks once.  The next step's slabs load into
// registers while the MFMAs of the current one run.
//
// TM = float is the mixed-precision former's variant: an FP32 A, B rounded to FP32, and
// v_mfma_f32_16x16x4f32 (twice the FP64 rate) accumulating in FP32 over one split's K range;
// the split-K partials are summed in FP64.  Its LDS rows are 16 floats apart modulo 64 banks.
// (64-deep slabs, 256 contiguous bytes of each A row per step, were slower: 3 workgroups per CU
// instead of 6.)
template <typename TA, int LP, bool VEC, typename TM = double>
__global__ __launch_bounds__(256) void tall_kernel(GemmF64Args p) {
  constexpr int BM = 128, BK = 32;
  constexpr int KT = BK / 2;                 // consecutive k of one A row per thread
  constexpr int SA = BM + 16, SB = LP + 16;  // LDS row strides (elements)
  constexpr int NB = BK * LP / 256;          // B elements per thread
  constexpr int VA = VEC ? 16 / static_cast<int>(sizeof(TA)) : 1;
  constexpr bool F32 = sizeof(TM) == 4;
  static_assert(!F32 || sizeof(TA) == 4, "FP32 MFMA needs an FP32 A");
  typedef typename VecOf<TA, VA>::type V;
  typedef typename std::conditional<F32, f4, d4>::type Acc;
  __shared__ TM sA[BK * SA];
  __shared__ TM sB[BK * SB];
  if (p.gate != nullptr && *p.gate < p.gate_t) return;
  const int64_t tiles_m = ceil_div(p.M, BM);
  const int64_t tile = blockIdx.x % tiles_m, split = blockIdx.x / tiles_m;
  const int64_t i0 = tile * BM;
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const TA* A = static_cast<const TA*>(p.A);
  const double* B = static_cast<const double*>(p.B);
  // A slab: thread -> row ra = tid / 2, k offset ka = KT * (tid % 2)
  const int ra = tid / 2, ka = KT * (tid % 2);
  const bool row_ok = i0 + ra < p.M;
  const TA* arow = A + (row_ok ? (i0 + ra) : 0) * p.lda;
  TM va[KT], vb[NB];
  auto load = [&](int64_t k0) {
    const int64_t kk = k0 + ka;
    bool done = false;
    if constexpr (VEC) {
      if (row_ok && kk + KT <= kend) {
#pragma unroll
        for (int v = 0; v < KT / VA; ++v) {
          const V x = *reinterpret_cast<const V*>(arow + kk + v * VA);
#pragma unroll
          for (int e = 0; e < VA; ++e) va[v * VA + e] = static_cast<TM>(x[e]);
        }
        done = true;
      }
    }
    if (!done) {
#pragma unroll
      for (int e = 0; e < KT; ++e)
        va[e] = (row_ok && kk + e < kend) ? static_cast<TM>(arow[kk + e]) : TM(0);
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = e * 256 + tid, kb = idx / LP, jb = idx % LP;
      vb[e] = (jb < p.N && k0 + kb < kend) ? static_cast<TM>(B[(k0 + kb) * p.ldb + jb]) : TM(0);
    }
  };
  auto store = [&] {
#pragma unroll
    for (int e = 0; e < KT; ++e) sA[(ka + e) * SA + ra] = va[e];
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = e * 256 + tid;
      sB[(idx / LP) * SB + idx % LP] = vb[e];
    }
  };
  constexpr int TJ = LP / 16;
  Acc acc[2][TJ];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b) acc[a][b] = Acc{0, 0, 0, 0};
  const int fr = lane / 16, fc = lane % 16;
  if (kbeg < kend) load(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();  // the previous step's fragment reads are done
    store();
    __syncthreads();
    if (k0 + BK < kend) load(k0 + BK);
#pragma unroll
    for (int s4 = 0; s4 < BK / 4; ++s4) {
      const int kr = 4 * s4 + fr;
      TM af[2], bf[TJ];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = sA[kr * SA + wave * 32 + a * 16 + fc];
#pragma unroll
      for (int b = 0; b < TJ; ++b) bf[b] = sB[kr * SB + b * 16 + fc];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b) {
          if constexpr (F32)
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
          else
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
        }
    }
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

// ------------------------------------------------------------------------------------------------
// Tall products with both operands streamed from HBM into LDS by DMA (global_load_lds_dwordx4)
// through an NS-stage ring, NS - 1 stages in flight: an FP32 A on FP32 MFMA (the mixed former's
// Z = A'Qbar of a column-major A, Qbar = A X of a row-major A) or an FP64 A on FP64 MFMA
// (Qbar = Y (T M), the FP64 former's Z, the Wen-Yin integrators' Bs products).  Rows of A run
// along K (element (i, k) at A[i*lda + k]); B is FP64 and row-major (element (k, j) at
// B[k*ldb + j]), rounded to FP32 as its fragments are read when A is FP32.  The register-staged tall_kernel keeps one step of
// A in flight per thread, which leaves this HBM-bound stream latency-bound (3.2 TB/s for the
// former at 100000 x 10000, l = 32); with three stages and two workgroups per CU this kernel
// streams it at 5.1 TB/s.  A itself reads at 6.1 TB/s with plain loads whatever the length of the
// row pieces (128 bytes to 1 KiB), and loading A and B straight into registers (no LDS) stays at
// 3.5 TB/s, MFMAs or not: 16 rows x 64 bytes per load instruction is slower to issue.
//
// A stage covers the 8 16-byte chunks (128 bytes) of each A row that hold BK = 8 E k (E = 4 FP32
// or 2 FP64 elements per chunk): a 128-row slab of A (16 KiB) and a BK x LP FP64 slab of B, in
// one LDS array per stage (distinct arrays addressed with static indices keep the compiler's DMA
// wait tracking per stage).  A DMA wave instruction writes 1 KiB of LDS contiguously, 16 bytes
// per lane; which 16-byte chunk of its rows a lane fetches is permuted so that the fragment reads
// below meet each of the 64 banks once per 16 lanes (b128) or 32 lanes (b64):
//   A row r, chunk c (k = E c .. E c + E - 1) at 16-byte unit 8 r + (c ^ ((r >> 1) & 7));
//   B row k, chunk c (j = 2c, 2c+1)           at unit (LP / 2) k + (c ^ (((k / E) & 1) << 3)).
// MFMA e (e < E) of chunk group q contracts lane L's k = E (4 q + L / 16) + e, so one b128 read
// gives a lane its E A values, and E b64 reads its E B values, of the same k.  Chunks past the
// split's K range are fetched from its first chunk and zeroed in the A fragments; rows past M or
// columns past N are fetched from row / column 0 and feed outputs that are never stored, so no
// DMA leaves the operands.  Needs K, lda multiples of E, N and ldb even, 16-byte aligned A and B.
// FP32 A: FP32 sums within one split's K range, FP64 across splits.
// LDS bytes of one stage: a 128-row A slab of 128 bytes per row, and its K x LP FP64 B slab.
template <typename TA, int LP>
constexpr int stream_stage_bytes() {
  return 128 * 128 + (128 / static_cast<int>(sizeof(TA))) * LP * 8;
}

template <typename TA, int LP, int NS>
__global__ __launch_bounds__(256, 163840 / (NS * stream_stage_bytes<TA, LP>()))
void tall_stream_kernel(GemmF64Args p) {
  constexpr int EPC = 16 / static_cast<int>(sizeof(TA));  // elements of A per 16-byte chunk
  constexpr bool F32 = sizeof(TA) == 4;
  constexpr int BM = 128, BK = 8 * EPC;  // a stage: 8 chunks (128 bytes) of each A row
  constexpr int TJ = LP / 16;
  constexpr int AU = BM * 8;           // 16-byte units of a stage's A slab (1024)
  constexpr int BU = BK * LP / 2;      // ... and of its B slab
  constexpr int LA = AU / 256, LB = BU / 256;  // DMA instructions per wave and stage
  constexpr int L = LA + LB;
  constexpr int SW = (AU + BU) * 4;  // floats per stage
  static_assert(BU % 256 == 0, "B slab in whole DMA instructions per wave");
  static_assert(NS >= 2 && NS <= 6, "stage count");
  static_assert((NS - 2) * L < 64, "vmcnt range");
  typedef typename std::conditional<F32, f4, d4>::type Acc;
  typedef typename VecOf<TA, EPC>::type V;
  __shared__ __attribute__((aligned(16))) float S0[SW];
  __shared__ __attribute__((aligned(16))) float S1[SW];
  __shared__ __attribute__((aligned(16))) float S2[NS > 2 ? SW : 4];
  __shared__ __attribute__((aligned(16))) float S3[NS > 3 ? SW : 4];
  __shared__ __attribute__((aligned(16))) float S4[NS > 4 ? SW : 4];
  __shared__ __attribute__((aligned(16))) float S5[NS > 5 ? SW : 4];
  if (p.gate != nullptr && *p.gate < p.gate_t) return;
  const int64_t tiles_m = ceil_div(p.M, BM);
  const int64_t tile = blockIdx.x % tiles_m, split = blockIdx.x / tiles_m;
  const int64_t i0 = tile * BM;
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;
  const int lane = threadIdx.x % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kWave));
  const TA* A = static_cast<const TA*>(p.A);
  const double* B = static_cast<const double*>(p.B);

  // Per-lane DMA sources, relative to the stage's first k.
  const TA* asrc[LA];
  int achunk[LA];
#pragma unroll
  for (int i = 0; i < LA; ++i) {
    const int r = 8 * (i * 4 + wave) + lane / 8;
    const int c = (lane % 8) ^ ((r >> 1) & 7);
    const int64_t row = (i0 + r < p.M) ? i0 + r : i0;
    asrc[i] = A + row * p.lda + EPC * c;
    achunk[i] = EPC * c;
  }
  constexpr int BR = 64 * 16 / (LP * 8);  // B rows per DMA instruction (4 or 2)
  const double* bsrc[LB];
  int brow[LB];
#pragma unroll
  for (int i = 0; i < LB; ++i) {
    const int k = BR * (i * 4 + wave) + lane / (LP / 2);
    int c = (lane % (LP / 2)) ^ (((k / EPC) & 1) << 3);
    if (2 * c >= p.N) c = 0;
    bsrc[i] = B + 2 * c;
    brow[i] = k;
  }
  // k within the split counted from kbeg in 32 bits (no 64-bit compares in the loop)
  const int krem = kbeg < kend ? static_cast<int>(kend - kbeg) : 0;
  const int ntk = (krem + BK - 1) / BK;
  // DMA instruction i (A for i < LA, then B) of stage kt; a chunk past the split's end reads
  // the split's first chunk
  auto dma = [&](int i, int kt, float* s) {
    const int kofs = kt * BK;
    if (i < LA) {
      dma16(asrc[i] + (kofs + achunk[i] < krem ? kbeg + kofs : kbeg - achunk[i]),
            (lds_void*)&s[(i * 4 + wave) * 256]);
    } else {
      const int q = i - LA;
      const int64_t k = kbeg + ((kofs + brow[q] < krem) ? kofs + brow[q] : 0);
      dma16(bsrc[q] + k * p.ldb, (lds_void*)&s[AU * 4 + (q * 4 + wave) * 256]);
    }
  };
  auto issue = [&](int kt, float* s) {
#pragma unroll
    for (int i = 0; i < L; ++i) dma(i, kt, s);
  };

  constexpr int TA2 = 2;  // 16-row groups per wave
  Acc acc[TA2][TJ];
#pragma unroll
  for (int a = 0; a < TA2; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b) acc[a][b] = Acc{0, 0, 0, 0};
  const int g = lane / 16, fc = lane % 16;
  const int bsw = (g & 1) << 3;
  // MASK: the stage may hold chunks past the split's range (its last stage and the padding
  // steps); whole stages run without the masks.  The stage NS - 1 ahead is issued half after each
  // chunk group's fragment reads, under the MFMAs (issued as one burst after the barrier, the DMA
  // held every wave's first MFMAs back).
  auto step = [&](auto mask, int kt, const float* cur, float* nxt) {
    constexpr bool MASK = decltype(mask)::value;
    asm volatile("" ::: "memory");
    wait_vmcnt<(NS - 2) * L>();    // this stage landed; the younger ones may still be in flight
    __builtin_amdgcn_s_barrier();  // ... for every wave; and the oldest buffer is fully read
    asm volatile("" ::: "memory");
    const int kn = kt + NS - 1 < ntk ? kt + NS - 1 : ntk - 1;
    const int kofs = kt * BK;
    const double* cb = reinterpret_cast<const double*>(cur + AU * 4);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = 4 * q + g;
      const bool ok = kofs + EPC * c < krem;
      V af[TA2];
#pragma unroll
      for (int a = 0; a < TA2; ++a) {
        const int r = wave * 32 + a * 16 + fc;
        af[a] = *reinterpret_cast<const V*>(&cur[(r * 8 + (c ^ ((r >> 1) & 7))) * 4]);
      }
      TA bf[TJ][EPC];
#pragma unroll
      for (int b = 0; b < TJ; ++b) {
        const int j = 16 * b + fc;
#pragma unroll
        for (int e = 0; e < EPC; ++e) {
          const int k = EPC * c + e;
          bf[b][e] = static_cast<TA>(cb[(k * (LP / 2) + ((j >> 1) ^ bsw)) * 2 + (j & 1)]);
        }
      }
#pragma unroll
      for (int i = q * L / 2; i < (q + 1) * L / 2; ++i) dma(i, kn, nxt);
#pragma unroll
      for (int e = 0; e < EPC; ++e)
#pragma unroll
        for (int a = 0; a < TA2; ++a)
#pragma unroll
          for (int b = 0; b < TJ; ++b) {
            const TA x = (!MASK || ok) ? af[a][e] : TA(0);
            if constexpr (F32)
              acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(x, bf[b][e], acc[a][b], 0, 0, 0);
            else
              acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, bf[b][e], acc[a][b], 0, 0, 0);
          }
    }
  };
  auto sbuf = [&](int s) -> float* {
    return s == 0 ? S0 : s == 1 ? S1 : s == 2 ? S2 : s == 3 ? S3 : s == 4 ? S4 : S5;
  };
  if (ntk > 0) {
#pragma unroll
    for (int s = 0; s < NS - 1; ++s) issue(s < ntk ? s : ntk - 1, sbuf(s));
    const int nsteps = (ntk + NS - 1) / NS * NS;
    const int nfull = (krem / BK) / NS * NS;  // whole stages, NS at a time
    const std::integral_constant<bool, false> unmasked;
    const std::integral_constant<bool, true> masked;
    int kt = 0;
    for (; kt < nfull; kt += NS) {
#pragma unroll
      for (int s = 0; s < NS; ++s) step(unmasked, kt + s, sbuf(s), sbuf((s + NS - 1) % NS));
    }
    for (; kt < nsteps; kt += NS) {
#pragma unroll
      for (int s = 0; s < NS; ++s) step(masked, kt + s, sbuf(s), sbuf((s + NS - 1) % NS));
    }
    wait_vmcnt<0>();  // no DMA may still target LDS when the workgroup ends
  }
  // D element r of a 16 x 16 tile sits at row 4 (lane/16) + r (FP32) or lane/16 + 4r (FP64),
  // column lane%16
  const bool to_partial = p.partial != nullptr;
  double* P = to_partial ? p.partial + split * p.M * p.N : nullptr;
#pragma unroll
  for (int a = 0; a < TA2; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = i0 + wave * 32 + a * 16 + (F32 ? 4 * g + r : g + 4 * r);
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

// The tall_direct_tile product (tall_direct.hpp) as a GEMM launch.
template <typename TA, typename TM, int LP, bool VEC>
__global__ __launch_bounds__(256, LP == 32 ? 3 : 2) void tall_direct_kernel(GemmF64Args p) {
  if (p.gate != nullptr && *p.gate < p.gate_t) return;
  tall_direct_tile<TA, TM, LP, VEC>(p, blockIdx.x);
}

template <typename TA, typename TM, int LP>
void launch_tall_direct(const GemmF64Args& p, hipStream_t st) {
  const int64_t nblocks = ceil_div(p.M, 128) * p.splits;
  MCNLA_REQUIRE(nblocks < (1LL << 31), MCNLA_ERR_INVALID, "GEMM grid too large");
  constexpr int ve = 16 / static_cast<int>(sizeof(TA));
  const bool vec = reinterpret_cast<uintptr_t>(p.A) % 16 == 0 && p.lda % ve == 0 &&
                   p.kchunk % ve == 0;
  if (vec)
    hipLaunchKernelGGL((tall_direct_kernel<TA, TM, LP, true>),
                       dim3(static_cast<unsigned>(nblocks)), dim3(256), 0, st, p);
  else
    hipLaunchKernelGGL((tall_direct_kernel<TA, TM, LP, false>),
                       dim3(static_cast<unsigned>(nblocks)), dim3(256), 0, st, p);
  MCNLA_LAUNCHED();
}

// Compute units of the current device (cached per device).
int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  MCNLA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) dev = 0;
  int n = cache[dev].load(std::memory_order_relaxed);
  if (n <= 0) {
    MCNLA_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    if (n <= 0) n = 1;
    cache[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

// A sizes (bytes) from which an FP32 A on FP32 MFMA streams by DMA by default.
constexpr double kStreamMinBytes = 1073741824.0;

// tall_stream_kernel's operand conditions (the others take the register-staged kernels).
bool tall_stream_ok(const GemmF64Args& p) {
  auto al16 = [](const void* q) { return reinterpret_cast<uintptr_t>(q) % 16 == 0; };
  const int64_t e = p.a_f32 ? 4 : 2;  // A elements per 16-byte chunk
  // The default kernel for an FP32 A of at least kStreamMinBytes on FP32 MFMA (smaller ones run
  // faster register-staged, alone and beside other solves), the alternative one otherwise.
  const double abytes = static_cast<double>(p.M) * static_cast<double>(p.K) * (p.a_f32 ? 4 : 8);
  const bool chosen = p.a_f32 ? p.f32_mfma && ((p.kernel == 0) == (abytes >= kStreamMinBytes))
                              : p.kernel == 1;
  return chosen && al16(p.A) && al16(p.B) && p.lda % e == 0 && p.K % e == 0 && p.K >= e &&
         p.N % 2 == 0 && p.ldb % 2 == 0;
}

// Stages of the streaming kernel's ring (FP32 A: 24 KiB each at N <= 32, 32 KiB at N <= 64;
// FP64 A: 20 and 24 KiB), two workgroups per CU in each case.
template <typename TA, int LP>
constexpr int stream_stages() {
  return sizeof(TA) == 4 ? (LP <= 32 ? 3 : 2) : (LP <= 32 ? 4 : 3);
}
template <typename TA, int LP>
constexpr int stream_per_cu() {
  return 163840 / (stream_stages<TA, LP>() * stream_stage_bytes<TA, LP>());
}
// A streaming kernel's grid is best a whole number of rounds of the workgroups the device runs
// at once: among the split counts from half the requested one up to it, take the one whose grid
// fills its last round best (the requested one on ties), with K ranges in whole 32-deep stages.
void tall_stream_plan(GemmF64Args& p, int64_t per_cu) {
  constexpr int64_t kmult = 32;
  const int64_t tiles = ceil_div(p.M, 128);
  const int64_t slots = static_cast<int64_t>(device_cus()) * per_cu;
  int64_t best = p.splits;
  double best_eff = 0.0;
  for (int64_t s = p.splits; s >= 1 && 2 * s >= p.splits; --s) {
    const int64_t w = tiles * s;
    const double eff = static_cast<double>(w) / static_cast<double>(ceil_div(w, slots) * slots);
    if (eff > best_eff + 1e-3) best_eff = eff, best = s;
  }
  p.kchunk = ceil_div(ceil_div(p.K, best), kmult) * kmult;
  p.splits = ceil_div(p.K, p.kchunk);
}

template <typename TA, int LP>
void launch_tall_stream(const GemmF64Args& p, hipStream_t st) {
  const int64_t nblocks = ceil_div(p.M, 128) * p.splits;
  MCNLA_REQUIRE(nblocks < (1LL << 31), MCNLA_ERR_INVALID, "GEMM grid too large");
  hipLaunchKernelGGL((tall_stream_kernel<TA, LP, stream_stages<TA, LP>()>),
                     dim3(static_cast<unsigned>(nblocks)), dim3(256), 0, st, p);
  MCNLA_LAUNCHED();
}

template <typename TA, int LP, typename TM = double>
void launch_tall(const GemmF64Args& p, hipStream_t st) {
  const int64_t nblocks = ceil_div(p.M, 128) * p.splits;
  MCNLA_REQUIRE(nblocks < (1LL << 31), MCNLA_ERR_INVALID, "GEMM grid too large");
  constexpr int va = 16 / static_cast<int>(sizeof(TA));
  const bool vec = reinterpret_cast<uintptr_t>(p.A) % 16 == 0 && p.lda % va == 0 &&
                   p.kchunk % va == 0;
  if (vec)
    hipLaunchKernelGGL((tall_kernel<TA, LP, true, TM>), dim3(static_cast<unsigned>(nblocks)),
                       dim3(256), 0, st, p);
  else
    hipLaunchKernelGGL((tall_kernel<TA, LP, false, TM>), dim3(static_cast<unsigned>(nblocks)),
                       dim3(256), 0, st, p);
  MCNLA_LAUNCHED();
}

// Deterministic split-K reduction: C = alpha * sum_s partial[s] + beta * C (sums in split order).
__global__ void splitk_reduce_kernel(const double* __restrict__ partial, int64_t splits,
                                     int64_t batch, int64_t M, int64_t N, double alpha,
                                     double beta, double* __restrict__ C, int64_t ldc,
                                     int64_t stride_c, const int* gate, int gate_t) {
  if (gate != nullptr && *gate < gate_t) return;
  const int64_t per = M * N;
  const int64_t total = per * batch;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / per, ij = e % per;
    const int64_t i = ij / N, j = ij % N;
    const double* pb = partial + b * splits * per + ij;
    // eight loads in flight per thread (one at a time leaves this stream latency-bound), summed
    // in split order
    double s = 0.0;
    int64_t q = 0;
    for (; q + 8 <= splits; q += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = pb[(q + u) * per];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; q < splits; ++q) s += pb[q * per];
    double* c = C + b * stride_c + i * ldc + j;
    *c = (beta == 0.0) ? alpha * s : alpha * s + beta * *c;
  }
}

// The reduction of a symmetric output (every entry with j <= i computed), by 64 x 64 tiles of the
// lower triangle: a tile's sums (split order, as above) go through LDS, so that both it and its
// mirror image above the diagonal are written in whole rows (a mirror written entry by entry
// touches one cache line per entry).  Grid: (T (T + 1) / 2 tiles, batch) of 1024 threads,
// T = ceil(N / 64).
__global__ __launch_bounds__(1024) void splitk_reduce_sym_kernel(
    const double* __restrict__ partial, int64_t splits, int64_t N, double alpha, double beta,
    double* __restrict__ C, int64_t ldc, int64_t stride_c, const int* gate, int gate_t) {
  __shared__ double t[64][65];
  if (gate != nullptr && *gate < gate_t) return;
  const int64_t x = blockIdx.x;
  int64_t ti = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(x) + 1.0) - 1.0) / 2.0);
  while (ti * (ti + 1) / 2 > x) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= x) ++ti;
  const int64_t tj = x - ti * (ti + 1) / 2;
  const int64_t per = N * N;
  const double* pb = partial + static_cast<int64_t>(blockIdx.y) * splits * per;
  double* Cb = C + static_cast<int64_t>(blockIdx.y) * stride_c;
  const int c = threadIdx.x % 64, r0 = threadIdx.x / 64;
  const bool diag = ti == tj;
  // the thread's 4 entries (rows r0, r0 + 16, ...), summed split by split: 4 splits x 4 entries
  // of loads in flight (and few registers, so that 8 waves share a SIMD), each entry's sum in
  // split order
  constexpr int RU = 4;
  int64_t off[RU];
  bool ok[RU];
  double acc[RU];
#pragma unroll
  for (int u = 0; u < RU; ++u) {
    const int64_t i = ti * 64 + r0 + 16 * u, j = tj * 64 + c;
    ok[u] = i < N && j < N && (!diag || j <= i);
    off[u] = ok[u] ? i * N + j : 0;
    acc[u] = 0.0;
  }
  int64_t q = 0;
  for (; q + 4 <= splits; q += 4) {
    double v[4][RU];
#pragma unroll
    for (int w = 0; w < 4; ++w)
#pragma unroll
      for (int u = 0; u < RU; ++u) v[w][u] = pb[(q + w) * per + off[u]];
#pragma unroll
    for (int w = 0; w < 4; ++w)
#pragma unroll
      for (int u = 0; u < RU; ++u) acc[u] += v[w][u];
  }
  for (; q < splits; ++q)
#pragma unroll
    for (int u = 0; u < RU; ++u) acc[u] += pb[q * per + off[u]];
#pragma unroll
  for (int u = 0; u < RU; ++u) t[r0 + 16 * u][c] = ok[u] ? acc[u] : 0.0;
  __syncthreads();
  for (int r = r0; r < 64; r += 16) {
    const int64_t i = ti * 64 + r, j = tj * 64 + c;
    if (i < N && j < N) {
      const double v = (!diag || j <= i) ? t[r][c] : t[c][r];
      double* d = Cb + i * ldc + j;
      *d = (beta == 0.0) ? alpha * v : alpha * v + beta * *d;
    }
    if (!diag) {
      const int64_t i2 = tj * 64 + r, j2 = ti * 64 + c;  // the mirror tile, row by row
      if (i2 < N && j2 < N) {
        const double v = t[c][r];
        double* d = Cb + i2 * ldc + j2;
        *d = (beta == 0.0) ? alpha * v : alpha * v + beta * *d;
      }
    }
  }
}

template <typename TA, typename TB, bool A_MC, bool B_NC, int BM, int BN, int BK, int WM, int WN>
void launch_cfg(const GemmF64Args& p, hipStream_t st) {
  using Cfg = TileCfg<BM, BN, BK, WM, WN>;
  const int64_t tiles_m = ceil_div(p.M, BM), tiles_n = ceil_div(p.N, BN);
  const int64_t ntiles = p.lower_only ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
  const int64_t nblocks = ntiles * p.splits * p.batch;
  MCNLA_REQUIRE(nblocks < (1LL << 31), MCNLA_ERR_INVALID, "GEMM grid too large");
  // 16-byte loads need 16-byte aligned operands whose leading dimensions and batch strides are
  // multiples of the vector length.
  auto aligned = [](const void* ptr, int64_t ld, int64_t stride, int64_t vec) {
    return (reinterpret_cast<uintptr_t>(ptr) % 16 == 0) && ld % vec == 0 && stride % vec == 0;
  };
  const bool vec = aligned(p.A, p.lda, p.stride_a, 16 / sizeof(TA)) &&
                   aligned(p.B, p.ldb, p.stride_b, 16 / sizeof(TB));
  if (vec)
    hipLaunchKernelGGL((gemm_f64_kernel<TA, TB, A_MC, B_NC, BM, BN, BK, WM, WN, true>),
                       dim3(static_cast<unsigned>(nblocks)), dim3(Cfg::kThreads), 0, st, p);
  else
    hipLaunchKernelGGL((gemm_f64_kernel<TA, TB, A_MC, B_NC, BM, BN, BK, WM, WN, false>),
                       dim3(static_cast<unsigned>(nblocks)), dim3(Cfg::kThreads), 0, st, p);
  MCNLA_LAUNCHED();
}

template <typename TA, typename TB, bool A_MC, bool B_NC>
void launch_shape(const GemmF64Args& p, hipStream_t st) {
  switch (p.shape) {
    case GemmShape::kSquare:  // 128 x 128 tiles, 4 waves of 64 x 64
      launch_cfg<TA, TB, A_MC, B_NC, 128, 128, 16, 64, 64>(p, st);
      break;
    case GemmShape::kTallN32:  // 128 x 32 tiles, 4 waves stacked along M
      launch_cfg<TA, TB, A_MC, B_NC, 128, 32, 16, 32, 32>(p, st);
      break;
    case GemmShape::kTallN64:
      launch_cfg<TA, TB, A_MC, B_NC, 128, 64, 16, 32, 64>(p, st);
      break;
    case GemmShape::kWideM32:  // 32 x 128 tiles, 4 waves along N
      launch_cfg<TA, TB, A_MC, B_NC, 32, 128, 16, 32, 32>(p, st);
      break;
    case GemmShape::kSmall:  // 32 x 32 tiles, 1 wave (tiny l x l products, split-K)
      launch_cfg<TA, TB, A_MC, B_NC, 32, 32, 16, 32, 32>(p, st);
      break;
  }
}

template <typename TA, typename TB>
void launch_types(const GemmF64Args& p, hipStream_t st) {
  if (p.a_mcontig) {
    if (p.b_ncontig) launch_shape<TA, TB, true, true>(p, st);
    else launch_shape<TA, TB, true, false>(p, st);
  } else {
    if (p.b_ncontig) launch_shape<TA, TB, false, true>(p, st);
    else launch_shape<TA, TB, false, false>(p, st);
  }
}

}  // namespace

void gemm_f64(GemmF64Args p, hipStream_t st) {
  MCNLA_REQUIRE(p.M >= 0 && p.N >= 0 && p.K >= 0 && p.batch >= 1, MCNLA_ERR_INVALID,
                "bad GEMM size");
  if (p.M == 0 || p.N == 0) return;
  if (p.splits < 1) p.splits = 1;
  if (p.kchunk <= 0) p.kchunk = ceil_div(p.K, p.splits);
  auto even16 = [](const void* ptr, int64_t ld, int64_t stride) {
    return reinterpret_cast<uintptr_t>(ptr) % 16 == 0 && ld % 2 == 0 && stride % 2 == 0;
  };
  // 256 x 128 DMA tiles from N = 128, 256 x 64 ones for 32 < N <= 64 (an even K: B's k-rows go
  // in pairs; no lower-triangle schedule)
  const bool narrow = p.N > 32 && p.N <= 64 && !p.lower_only && p.K % 2 == 0;
  const bool dma = p.kernel == 0 && !p.a_f32 && !p.b_f32 && p.a_mcontig && p.b_ncontig &&
                   p.shape == GemmShape::kSquare && p.M >= 128 && (p.N >= 128 || narrow) &&
                   even16(p.A, p.lda, p.stride_a) &&
                   even16(p.B, p.ldb, p.stride_b);
  const bool tall = !p.a_mcontig && p.b_ncontig && !p.b_f32 && p.N <= 64 && p.batch == 1 &&
                    !p.lower_only && p.M >= 128 &&
                    (p.shape == GemmShape::kTallN32 || p.shape == GemmShape::kTallN64);
  const bool stream = tall && (p.splits == 1 || p.partial != nullptr) && tall_stream_ok(p);
  // The register-staged kernels need split ranges that are whole K tiles (16); the DMA kernels
  // mask a partial last tile, so their splits can be equal.
  if (stream) {
    if (p.a_f32)
      tall_stream_plan(p, p.N <= 32 ? stream_per_cu<float, 32>() : stream_per_cu<float, 64>());
    else
      tall_stream_plan(p, p.N <= 32 ? stream_per_cu<double, 32>() : stream_per_cu<double, 64>());
  }
  else if (!dma)
    p.kchunk = ceil_div(p.kchunk, 16) * 16;
  if (p.K > 0) p.splits = ceil_div(p.K, p.kchunk);
  if (p.splits == 1 && !p.lower_only && p.C != nullptr) p.partial = nullptr;
  MCNLA_REQUIRE(!(p.splits > 1 && p.partial == nullptr), MCNLA_ERR_INVALID,
                "split-K GEMM needs a partial buffer");
  MCNLA_REQUIRE(!p.lower_only || p.M == p.N, MCNLA_ERR_INVALID, "lower_only needs M == N");
  MCNLA_REQUIRE(!p.lower_only || p.shape == GemmShape::kSquare || p.shape == GemmShape::kSmall,
                MCNLA_ERR_INVALID, "lower_only needs square tiles");
  if (p.ev_begin) MCNLA_HIP(hipEventRecord(p.ev_begin, st));
  if (tall) {
    // defaults (the fastest measured at the benchmark shapes): both operands streamed into LDS
    // by DMA for an FP32 A of 1 GiB or more on FP32 MFMA (where the alignment conditions hold;
    // 0.79 against 1.25 ms for the 4 GB former Z = A'Qbar at 100000 x 10000), A staged through
    // registers for smaller FP32 A (serving-size requests run faster so, alone and in the mix)
    // and for an FP32 A in FP64 arithmetic, and A straight into registers for an FP64 A.
    // p.kernel = 1 takes the other one for an FP32 A on FP32 MFMA, A straight into registers
    // for an FP32 A in FP64 arithmetic, and the DMA stream for an FP64 A (the integrators' Bs
    // products ask for it: 174 against 196 us for B_s B at 8192 x 8192, l = 64; slower for
    // Qbar = Y (T M): 0.42 against 0.40 ms at 100000 x 2048).
    if (p.a_f32 && p.f32_mfma) {
      if (stream) {
        if (p.N <= 32) launch_tall_stream<float, 32>(p, st);
        else launch_tall_stream<float, 64>(p, st);
      } else {
        if (p.N <= 32) launch_tall<float, 32, float>(p, st);
        else launch_tall<float, 64, float>(p, st);
      }
    } else if (p.a_f32) {
      if (p.kernel == 1) {
        if (p.N <= 32) launch_tall_direct<float, double, 32>(p, st);
        else launch_tall_direct<float, double, 64>(p, st);
      } else {
        if (p.N <= 32) launch_tall<float, 32>(p, st);
        else launch_tall<float, 64>(p, st);
      }
    } else {
      if (stream) {
        if (p.N <= 32) launch_tall_stream<double, 32>(p, st);
        else launch_tall_stream<double, 64>(p, st);
      } else {
        if (p.N <= 32) launch_tall_direct<double, double, 32>(p, st);
        else launch_tall_direct<double, double, 64>(p, st);
      }
    }
  } else if (dma) {
    // 256 x 128 tiles of 8 waves (64 x 64 each); a lower-triangle schedule takes the tiles that
    // touch the diagonal (2 (tm + 1) of them in tile row tm).
    const int64_t tm = ceil_div(p.M, 256);
    if (p.N <= 64) {  // 256 x 64 tiles of 8 waves (64 x 32 each)
      const int64_t nblocks = tm * p.splits * p.batch;
      MCNLA_REQUIRE(nblocks < (1LL << 31), MCNLA_ERR_INVALID, "GEMM grid too large");
      hipLaunchKernelGGL((gemm_f64_dma_kernel<256, 64, 16, 64, 32>),
                         dim3(static_cast<unsigned>(nblocks)), dim3(512), 0, st, p);
    } else {
      const int64_t tiles = p.lower_only ? tm * tm + tm : tm * ceil_div(p.N, 128);
      const int64_t nblocks = tiles * p.splits * p.batch;
      MCNLA_REQUIRE(nblocks < (1LL << 31), MCNLA_ERR_INVALID, "GEMM grid too large");
      hipLaunchKernelGGL((gemm_f64_dma_kernel<256, 128, 16, 64, 64>),
                         dim3(static_cast<unsigned>(nblocks)), dim3(512), 0, st, p);
    }
    MCNLA_LAUNCHED();
  } else if (p.a_f32) {
    if (p.b_f32) launch_types<float, float>(p, st);
    else launch_types<float, double>(p, st);
  } else {
    if (p.b_f32) launch_types<double, float>(p, st);
    else launch_types<double, double>(p, st);
  }
  if (p.ev_end) MCNLA_HIP(hipEventRecord(p.ev_end, st));
  if (p.partial != nullptr && p.C != nullptr && p.lower_only) {
    const int64_t T = ceil_div(p.N, 64);
    MCNLA_REQUIRE(T * (T + 1) / 2 < (1LL << 31) && p.batch < 65536, MCNLA_ERR_INVALID,
                  "GEMM grid too large");
    hipLaunchKernelGGL(splitk_reduce_sym_kernel,
                       dim3(static_cast<unsigned>(T * (T + 1) / 2), static_cast<unsigned>(p.batch)),
                       dim3(1024), 0, st, p.partial, p.splits, p.N, p.alpha, p.beta, p.C, p.ldc,
                       p.stride_c, p.gate, p.gate_t);
    MCNLA_LAUNCHED();
  } else if (p.partial != nullptr && p.C != nullptr) {
    const int64_t total = p.M * p.N * p.batch;
    const int threads = 256;
    int64_t blocks = ceil_div(total, threads);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0,
                       st, p.partial, p.splits, p.batch, p.M, p.N, p.alpha, p.beta, p.C, p.ldc,
                       p.stride_c, p.gate, p.gate_t);
    MCNLA_LAUNCHED();
  }
}

size_t gemm_f64_partial_elems(int64_t M, int64_t N, int64_t splits, int64_t batch) {
  return static_cast<size_t>(M) * static_cast<size_t>(N) * static_cast<size_t>(splits) *
         static_cast<size_t>(batch);
}

int compute_units() { return device_cus(); }

int64_t gemm_f64_dma_tail(int64_t M, int64_t N, int64_t K, int64_t* tail_splits) {
  // 256 x 128 tiles, one workgroup per CU.  The rows of whole rounds run as they are; the rest
  // (less than half a round) split K so that their workgroups end sooner.
  *tail_splits = 1;
  const int64_t cus = device_cus(), tn = ceil_div(N, 128), tm = ceil_div(M, 256);
  int64_t g = cus, b = tn;  // row tiles per whole round: cus / gcd(cus, tn)
  while (b != 0) {
    const int64_t r = g % b;
    g = b;
    b = r;
  }
  const int64_t per_round = cus / g;
  const int64_t main_tm = tm / per_round * per_round;
  const int64_t tail = (tm - main_tm) * tn;  // workgroups of the last, partial round
  if (main_tm == 0 || tail == 0 || 2 * tail > cus || (M - main_tm * 256) < 128) return M;
  int64_t s = cus / tail;
  if (s > 4) s = 4;
  if (K / s < 1024) s = K / 1024;  // keep each split long
  if (s < 2) return M;
  *tail_splits = s;
  return main_tm * 256;
}

int64_t gemm_f64_splits_for(int64_t tiles, int64_t K, int64_t target_blocks) {
  int64_t s = ceil_div(target_blocks, tiles > 0 ? tiles : 1);
  const int64_t cap = ceil_div(K, 256);  // keep at least ~256 k per split
  if (s > cap) s = cap;
  if (s < 1) s = 1;
  return s;
}

}  // namespace mcnla
