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
#include "common.hpp"
#include "kernels.hpp"

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

// Loads one operand tile (rows = BK k-values, cols = MN values) from global memory into
// registers.  MN_CONTIG: element (k, mn) lives at base[mn + k*ld]; otherwise at base[mn*ld + k].
template <typename TS, bool MN_CONTIG, int MN, int BK, int NT, int ELEMS>
__device__ __forceinline__ void load_tile(double (&r)[ELEMS], const TS* __restrict__ base,
                                          int64_t ld, int64_t mn0, int64_t k0, int64_t mn_lim,
                                          int64_t k_lim, bool full) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int e = 0; e < ELEMS; ++e) {
    const int idx = e * NT + tid;
    int kk, mm;
    if (MN_CONTIG) {
      kk = idx / MN;
      mm = idx % MN;
    } else {
      mm = idx / BK;
      kk = idx % BK;
    }
    const int64_t gm = mn0 + mm, gk = k0 + kk;
    const int64_t off = MN_CONTIG ? (gm + gk * ld) : (gm * ld + gk);
    if (full || (gm < mn_lim && gk < k_lim)) {
      r[e] = static_cast<double>(base[off]);
    } else {
      r[e] = 0.0;
    }
  }
}

template <bool MN_CONTIG, int MN, int BK, int NT, int ELEMS, int LD>
__device__ __forceinline__ void store_tile(double* __restrict__ lds, const double (&r)[ELEMS]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int e = 0; e < ELEMS; ++e) {
    const int idx = e * NT + tid;
    int kk, mm;
    if (MN_CONTIG) {
      kk = idx / MN;
      mm = idx % MN;
    } else {
      mm = idx / BK;
      kk = idx % BK;
    }
    lds[kk * LD + mm] = r[e];
  }
}

template <typename TA, typename TB, bool A_MC, bool B_NC, int BM, int BN, int BK, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_f64_kernel(
    GemmF64Args p) {
  using Cfg = TileCfg<BM, BN, BK, WM, WN>;
  constexpr int NT = Cfg::kThreads;
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
      load_tile<TA, A_MC, BM, BK, NT, Cfg::kElemsA>(ra, A, p.lda, m0, kbeg, p.M, kend, full);
      load_tile<TB, B_NC, BN, BK, NT, Cfg::kElemsB>(rb, B, p.ldb, n0, kbeg, p.N, kend, full);
      store_tile<A_MC, BM, BK, NT, Cfg::kElemsA, Cfg::kLdA>(As[0], ra);
      store_tile<B_NC, BN, BK, NT, Cfg::kElemsB, Cfg::kLdB>(Bs[0], rb);
    }
    __syncthreads();
    for (int64_t kt = 0; kt < ntk; ++kt) {
      const int cur = kt & 1;
      const bool more = kt + 1 < ntk;
      if (more) {
        const int64_t k0 = kbeg + (kt + 1) * BK;
        const bool full = full_mn && (k0 + BK <= kend);
        load_tile<TA, A_MC, BM, BK, NT, Cfg::kElemsA>(ra, A, p.lda, m0, k0, p.M, kend, full);
        load_tile<TB, B_NC, BN, BK, NT, Cfg::kElemsB>(rb, B, p.ldb, n0, k0, p.N, kend, full);
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
        store_tile<A_MC, BM, BK, NT, Cfg::kElemsA, Cfg::kLdA>(As[cur ^ 1], ra);
        store_tile<B_NC, BN, BK, NT, Cfg::kElemsB, Cfg::kLdB>(Bs[cur ^ 1], rb);
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

// Deterministic split-K reduction: C = alpha * sum_s partial[s] + beta * C.  With sym_tile > 0
// only tiles (tm >= tn) of size sym_tile were computed; the upper triangle is mirrored.
__global__ void splitk_reduce_kernel(const double* __restrict__ partial, int64_t splits,
                                     int64_t batch, int64_t M, int64_t N, double alpha,
                                     double beta, double* __restrict__ C, int64_t ldc,
                                     int64_t stride_c, int sym_tile, const int* gate, int gate_t) {
  if (gate != nullptr && *gate < gate_t) return;
  const int64_t per = M * N;
  const int64_t total = per * batch;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / per, ij = e % per;
    const int64_t i = ij / N, j = ij % N;
    int64_t si = i, sj = j;
    if (sym_tile > 0 && (i / sym_tile) < (j / sym_tile)) {
      si = j;
      sj = i;
    }
    const double* pb = partial + b * splits * per;
    double s = 0.0;
    for (int64_t q = 0; q < splits; ++q) s += pb[q * per + si * N + sj];
    double* c = C + b * stride_c + i * ldc + j;
    *c = (beta == 0.0) ? alpha * s : alpha * s + beta * *c;
  }
}

template <typename TA, typename TB, bool A_MC, bool B_NC, int BM, int BN, int BK, int WM, int WN>
void launch_cfg(const GemmF64Args& p, hipStream_t st) {
  using Cfg = TileCfg<BM, BN, BK, WM, WN>;
  const int64_t tiles_m = ceil_div(p.M, BM), tiles_n = ceil_div(p.N, BN);
  const int64_t ntiles = p.lower_only ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
  const int64_t nblocks = ntiles * p.splits * p.batch;
  MCNLA_REQUIRE(nblocks < (1LL << 31), MCNLA_ERR_INVALID, "GEMM grid too large");
  hipLaunchKernelGGL((gemm_f64_kernel<TA, TB, A_MC, B_NC, BM, BN, BK, WM, WN>),
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
  // kchunk must be a multiple of the K tile (16) so split ranges tile exactly.
  p.kchunk = ceil_div(p.kchunk, 16) * 16;
  if (p.K > 0) p.splits = ceil_div(p.K, p.kchunk);
  if (p.splits == 1 && !p.lower_only && p.C != nullptr) p.partial = nullptr;
  MCNLA_REQUIRE(!(p.splits > 1 && p.partial == nullptr), MCNLA_ERR_INVALID,
                "split-K GEMM needs a partial buffer");
  MCNLA_REQUIRE(!p.lower_only || p.M == p.N, MCNLA_ERR_INVALID, "lower_only needs M == N");
  MCNLA_REQUIRE(!p.lower_only || p.shape == GemmShape::kSquare || p.shape == GemmShape::kSmall,
                MCNLA_ERR_INVALID, "lower_only needs square tiles");
  if (p.a_f32) {
    if (p.b_f32) launch_types<float, float>(p, st);
    else launch_types<float, double>(p, st);
  } else {
    if (p.b_f32) launch_types<double, float>(p, st);
    else launch_types<double, double>(p, st);
  }
  if (p.partial != nullptr && p.C != nullptr) {
    const int sym = p.lower_only ? (p.shape == GemmShape::kSquare ? 128 : 32) : 0;
    const int64_t total = p.M * p.N * p.batch;
    const int threads = 256;
    int64_t blocks = ceil_div(total, threads);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0,
                       st, p.partial, p.splits, p.batch, p.M, p.N, p.alpha, p.beta, p.C, p.ldc,
                       p.stride_c, sym, p.gate, p.gate_t);
    MCNLA_LAUNCHED();
  }
}

size_t gemm_f64_partial_elems(int64_t M, int64_t N, int64_t splits, int64_t batch) {
  return static_cast<size_t>(M) * static_cast<size_t>(N) * static_cast<size_t>(splits) *
         static_cast<size_t>(batch);
}

int64_t gemm_f64_splits_for(int64_t tiles, int64_t K, int64_t target_blocks) {
  int64_t s = ceil_div(target_blocks, tiles > 0 ? tiles : 1);
  const int64_t cap = ceil_div(K, 256);  // keep at least ~256 k per split
  if (s > cap) s = cap;
  if (s < 1) s = 1;
  return s;
}

}  // namespace mcnla
