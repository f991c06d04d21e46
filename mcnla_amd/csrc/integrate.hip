// Integrators on the Gram matrix Bs = Qs' Qs (Nl x Nl), FP64.
//
// Both KN and Wen-Yin only touch Qs through products Qs'(Qs X) = Bs X, so once Bs is formed
// every iteration works on Nl x l and l x l matrices that stay in L2 (the reference's Gramian
// forms, row_block_gramian_kolmogorov_nagumo_integrator.hpp:106-226 and
// row_block_gramian_wen_yin_integrator.hpp:114-328).  The basis is carried as coefficients:
// Qc = Q0 F~ + Qs E~, so Qbar = Qs * M with M = E~ + [F~; 0].
//
// Every iteration t is a fixed sequence of kernels gated by st->converged_at: a kernel of
// iteration t does nothing once convergence was detected at an earlier iteration, so the host can
// enqueue a chunk of iterations and synchronize once per chunk.  Iteration t reads the buffers of
// parity t&1 and writes parity (t+1)&1.
#include "common.hpp"
#include "kernels.hpp"
#include "small.hpp"
#include "tall_direct.hpp"

#include <algorithm>
#include <climits>
#include <vector>

namespace mcnla {

struct IntegratorState {
  int converged_at;  // iteration index that met the stopping test, INT_MAX while running
  int nan_seen;      // a non-finite value appeared in the l x l step
  int ticket;        // workgroups done with the current fused reduction (reset by the last one)
  int pad;
  double res;        // last residual (KN: ||s - 1||_2, WY: mu)
  double taug, zeta, phi, mu;  // Wen-Yin scalars
};

namespace {

constexpr int kThreads = 256;
constexpr int kRows = 32;   // Nl rows per workgroup in the WY step-size pass
constexpr int kURows = 32;  // Nl rows per workgroup in the WY update
constexpr int kPRows = 64;  // Nl rows per workgroup in the partial dot products

__device__ __forceinline__ bool skip(const IntegratorState* st, int t) {
  return st->converged_at < t;
}

// Per-block partial products over kPRows rows of the Nl x l matrices B and Bg:
// P[blk][0] = B_blk' B_blk, P[blk][1] = B_blk' Bg_blk, P[blk][2] = Bg_blk' Bg_blk (nparts 3).
__global__ __launch_bounds__(kThreads) void gram_partials_kernel(int t, const IntegratorState* st,
                                                                int64_t nl, int l,
                                                                const double* __restrict__ B,
                                                                const double* __restrict__ Bg,
                                                                double* __restrict__ P,
                                                                int nparts) {
  if (skip(st, t)) return;
  __shared__ double sB[kPRows][kMaxL + 1];
  __shared__ double sG[kPRows][kMaxL + 1];
  const int64_t r0 = blockIdx.x * (int64_t)kPRows;
  for (int e = threadIdx.x; e < kPRows * l; e += kThreads) {
    const int rr = e / l, j = e % l;
    const int64_t r = r0 + rr;
    sB[rr][j] = (r < nl) ? B[r * l + j] : 0.0;
    sG[rr][j] = (r < nl) ? Bg[r * l + j] : 0.0;
  }
  __syncthreads();
  // the nparts l x l products over the block's rows on f64 MFMA tiles (16 x 16, k = row)
  double* Pb = P + blockIdx.x * (int64_t)nparts * l * l;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int nt1 = pad16(l) / 16, tiles = nt1 * nt1;
  for (int tile = wave; tile < nparts * tiles; tile += kThreads / kWave) {
    const int part = tile / tiles, i0 = ((tile % tiles) / nt1) * 16, j0 = (tile % nt1) * 16;
    const double(*X)[kMaxL + 1] = (part == 2) ? sG : sB;
    const double(*Y)[kMaxL + 1] = (part == 0) ? sB : sG;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < kPRows; k0 += 4) {
      const int k = k0 + lane / 16;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(X[k][i0 + lane % 16], Y[k][j0 + lane % 16], acc,
                                                 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + lane / 16 + 4 * r, j = j0 + lane % 16;
      if (i < l && j < l) Pb[part * l * l + i * l + j] = acc[r];
    }
  }
}

// sum_blk P[blk][q] in block order (deterministic).  The loads of 16 blocks are issued before
// any is added, so a thread waits for one memory round trip per 16 blocks, not per block.
__device__ __forceinline__ double sum_partials(const double* __restrict__ P, int64_t nblk,
                                               int64_t count, int64_t q) {
  double total = 0.0;
  for (int64_t b0 = 0; b0 < nblk; b0 += 16) {
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = (b0 + j < nblk) ? P[(b0 + j) * count + q] : 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j) total += v[j];
  }
  return total;
}

// Two such sums at once (32 loads in flight); q1 < 0 skips the second.
__device__ __forceinline__ void sum_partials2(const double* __restrict__ P, int64_t nblk,
                                              int64_t count, int64_t q0, int64_t q1, double& s0,
                                              double& s1) {
  s0 = 0.0;
  s1 = 0.0;
  for (int64_t b0 = 0; b0 < nblk; b0 += 16) {
    double v[16], w[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      v[j] = (b0 + j < nblk) ? P[(b0 + j) * count + q0] : 0.0;
      w[j] = (b0 + j < nblk && q1 >= 0) ? P[(b0 + j) * count + q1] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      s0 += v[j];
      s1 += w[j];
    }
  }
}

__global__ void reduce_partials_kernel(int t, const IntegratorState* st,
                                       const double* __restrict__ P, int64_t nblk, int count,
                                       double scale, double* __restrict__ out) {
  if (skip(st, t)) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  out[q] = scale * sum_partials(P, nblk, count, q);
}

// The KN l x l step (row_block_kolmogorov_nagumo_integrator.hpp:146-201): Dc = Bc'Bc/N and
// Dgc = Bc'Bgc/N from the per-block partials P (nblk blocks of [Bc'Bc, Bc'Bgc]), Z = Dgc - Dc^2,
// eig(Z), s = sqrt(1/2 + sqrt(1/4 - lambda)), C = V diag(s) V', Cinv = V diag(1/s) V',
// F = C - Dc Cinv, convergence when !(||s - 1||_2 >= tol).
//
// C and Cinv are matrix functions of Z, so they are computed here without the eigenvectors:
// X = (I - 4Z)^{1/2} and then C = ((I + X)/2)^{1/2}, Cinv = C^{-1} by two coupled Newton-Schulz
// iterations on MFMA tiles in LDS (both converge quadratically because Z is positive
// semidefinite with lambda <= 1/4), and ||s - 1||_2 = ||C - I||_F.  Should an iteration not
// converge (lambda at or above 1/4, non-finite input) the step falls back to a Jacobi
// eigen-decomposition and the formulas above, which also reproduce the reference's NaN there.
// Runs on NT threads (all of the workgroup).
template <int NT>
__device__ void kn_step_eigen(int t, IntegratorState* st, int l, double invN, double tol,
                              const double* __restrict__ P, int64_t nblk, double* __restrict__ F,
                              double* __restrict__ Cinv, double* dc, double* a, double* v,
                              double* z) {
  __shared__ double s[kMaxL];
  __shared__ double cs[2 * kMaxL];
  __shared__ int pq[2 * kMaxL];
  __shared__ double red[8];
  const int ld = lds_ld(l);
  const int ll = l * l;
  for (int e = threadIdx.x; e < 2 * ll; e += NT) {  // Dc into dc, Dgc into a
    const double x = invN * sum_partials(P, nblk, 2 * ll, e);
    const int q = e % ll;
    (e < ll ? dc : a)[(q / l) * ld + q % l] = x;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ll; e += NT) {  // Z = sym(Dgc) - Dc Dc'
    const int i = e / l, j = e % l;
    double x = 0.5 * (a[i * ld + j] + a[j * ld + i]);
    for (int q = 0; q < l; ++q) x -= dc[i * ld + q] * dc[j * ld + q];
    z[i * ld + j] = x;
  }
  __syncthreads();
  lds_jacobi(l, z, v, true, cs, pq, red);
  for (int i = threadIdx.x; i < l; i += NT) s[i] = sqrt(0.5 + sqrt(0.25 - z[i * ld + i]));
  __syncthreads();
  for (int e = threadIdx.x; e < ll; e += NT) {  // Cinv into `a`
    const int i = e / l, j = e % l;
    double ci = 0.0;
    for (int q = 0; q < l; ++q) ci += v[i * ld + q] * v[j * ld + q] / s[q];
    a[i * ld + j] = ci;
    Cinv[e] = ci;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ll; e += NT) {  // F = C - Dc Cinv, C on the fly
    const int i = e / l, j = e % l;
    double c = 0.0;
    for (int q = 0; q < l; ++q) c += v[i * ld + q] * v[j * ld + q] * s[q];
    double dci = 0.0;
    for (int q = 0; q < l; ++q) dci += dc[i * ld + q] * a[q * ld + j];
    F[e] = c - dci;
  }
  if (threadIdx.x == 0) {
    double r = 0.0;
    bool bad = false;
    for (int q = 0; q < l; ++q) {
      r += (s[q] - 1.0) * (s[q] - 1.0);
      if (!isfinite(s[q])) bad = true;
    }
    r = sqrt(r);
    st->res = r;
    if (bad) st->nan_seen = 1;
    if (!(r >= tol)) st->converged_at = t;
  }
}

// The NS form of the step; smem holds four (kMaxL x (kMaxL+1)) buffers, red NT/64 doubles.
template <int NT>
__device__ void kn_step_ns(int t, IntegratorState* st, int l, double invN, double tol,
                           const double* __restrict__ P, int64_t nblk, double* __restrict__ F,
                           double* __restrict__ Cinv, double* smem, double* red) {
  constexpr int NW = NT / kWave;
  constexpr int kBuf = kMaxL * (kMaxL + 1);
  double* sDc = smem;
  double* sY = smem + kBuf;
  double* sW = smem + 2 * kBuf;
  double* sT = smem + 3 * kBuf;
  const int np = pad16(l), ld = np + 1, ll = l * l;
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid & (kWave - 1);
  const int nt1 = np / 16, ntiles = nt1 * nt1;
  // Dc into sDc, Dgc into sT, both zero-padded to np x np
  for (int e = tid; e < np * np; e += NT) {
    const int i = e / np, j = e % np;
    double dc = 0.0, dgc = 0.0;
    if (i < l && j < l) sum_partials2(P, nblk, 2 * ll, i * l + j, ll + i * l + j, dc, dgc);
    sDc[i * ld + j] = invN * dc;
    sT[i * ld + j] = invN * dgc;
  }
  __syncthreads();
  // Y = I - 4Z with Z = sym(Dgc) - Dc Dc'
  for (int tile = wave; tile < ntiles; tile += NW) {
    const int i0 = (tile / nt1) * 16, j0 = (tile % nt1) * 16;
    const d4 acc = lds_mfma_tile<false, true>(np, ld, sDc, sDc, i0, j0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + (lane >> 4) + 4 * r, j = j0 + (lane & 15);
      const double zz = 0.5 * (sT[i * ld + j] + sT[j * ld + i]) - acc[r];
      sY[i * ld + j] = ((i == j) ? 1.0 : 0.0) - 4.0 * zz;
    }
  }
  __syncthreads();
  int ok = lds_ns_sqrt<NW>(np, ld, sY, sW, sT, red, 60);
  if (ok > 0) {
    for (int e = tid; e < np * np; e += NT) {
      const int i = e / np, j = e % np;
      sY[i * ld + j] = 0.5 * (((i == j) ? 1.0 : 0.0) + sY[i * ld + j]);
    }
    __syncthreads();
    ok = lds_ns_sqrt<NW>(np, ld, sY, sW, sT, red, 60);
  }
  if (ok <= 0) {  // rare: eigen-decomposition instead (uses its own layout of the buffers)
    kn_step_eigen<NT>(t, st, l, invN, tol, P, nblk, F, Cinv, sDc, sT, sW, sY);
    return;
  }
  // C = sY, Cinv = sW; F = C - Dc Cinv; ||C - I||_F
  double rs = 0.0;
  for (int tile = wave; tile < ntiles; tile += NW) {
    const int i0 = (tile / nt1) * 16, j0 = (tile % nt1) * 16;
    const d4 acc = lds_mfma_tile<false, false>(np, ld, sDc, sW, i0, j0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + (lane >> 4) + 4 * r, j = j0 + (lane & 15);
      if (i < l && j < l) {
        const double c = sY[i * ld + j];
        F[i * l + j] = c - acc[r];
        Cinv[i * l + j] = sW[i * ld + j];
        const double d = c - ((i == j) ? 1.0 : 0.0);
        rs += d * d;
      }
    }
  }
  rs = wave_sum(rs);
  __syncthreads();  // red is reused
  if (lane == 0) red[wave] = rs;
  __syncthreads();
  if (tid == 0) {
    double tot = 0.0;
    for (int w = 0; w < NW; ++w) tot += red[w];
    const double r = sqrt(tot);
    st->res = r;
    if (!isfinite(r)) st->nan_seen = 1;
    if (!(r >= tol)) st->converged_at = t;
  }
}

// One KN iteration's reductions and l x l step in one launch (after the split-K GEMM that
// leaves Pbg[s] = partial Bs * Bc over the s-th K range):
//   1. workgroup b owns rows [64 b, 64 b + 64) of the Nl x l matrices: Bgc = invN sum_s Pbg[s]
//      (fixed order), written out, and its block partials Bc'Bc, Bc'Bgc -> P[b] (f64 MFMA);
//   2. the last workgroup to finish (a ticket counter, taken after a device-scope fence; no
//      waiting) sums P over blocks in block order and runs the NS step above.
// The result is deterministic: no sum depends on which workgroup finishes last.
constexpr int kStepThreads = 512;
constexpr int kStepRows = 64;

// Workgroup `blk` of nblk (kStepThreads threads).
__device__ __forceinline__ void kn_reduce_step_body(
    int t, IntegratorState* st, int64_t nl, int l, double invN, double tol,
    const double* __restrict__ Pbg, int64_t splits, const double* __restrict__ B,
    double* __restrict__ Bg, double* __restrict__ P, int64_t nblk, double* __restrict__ F,
    double* __restrict__ Cinv, int64_t blk) {
  if (skip(st, t)) return;
  constexpr int kBuf = kMaxL * (kMaxL + 1);
  static_assert(2 * kStepRows * (kMaxL + 1) <= 4 * kBuf, "row blocks fit the step buffers");
  __shared__ double smem[4 * kBuf];
  __shared__ double red[kStepThreads / kWave];
  __shared__ int last;
  const int tid = threadIdx.x;
  const int ldr = kMaxL + 1;
  double* sB = smem;
  double* sG = smem + kStepRows * ldr;
  const int64_t r0 = blk * kStepRows;
  const int64_t per = nl * l;
  // elements e and e + kStepThreads together, so two sums' loads are in flight at once
  for (int e = tid; e < kStepRows * l; e += 2 * kStepThreads) {
    const int e2 = e + kStepThreads;
    const int64_t ra = r0 + e / l, rb = r0 + e2 / l;
    const bool va = ra < nl, vb = e2 < kStepRows * l && rb < nl;
    const int64_t oa = ra * l + e % l, ob = rb * l + e2 % l;
    double ga = 0.0, gb = 0.0;
    if (va) sum_partials2(Pbg, splits, per, oa, vb ? ob : -1, ga, gb);
    if (va) {
      ga *= invN;
      Bg[oa] = ga;
    }
    sB[(e / l) * ldr + e % l] = va ? B[oa] : 0.0;
    sG[(e / l) * ldr + e % l] = va ? ga : 0.0;
    if (e2 < kStepRows * l) {
      if (vb) {
        gb *= invN;
        Bg[ob] = gb;
      }
      sB[(e2 / l) * ldr + e2 % l] = vb ? B[ob] : 0.0;
      sG[(e2 / l) * ldr + e2 % l] = vb ? gb : 0.0;
    }
  }
  __syncthreads();
  const int ll = l * l;
  double* Pb = P + blk * (2 * ll);
  {  // [B'B, B'G] over the block's rows on f64 MFMA tiles: 2 (np/16)^2 tiles over the waves
    constexpr int NW = kStepThreads / kWave;
    const int wave = tid / kWave, lane = tid % kWave;
    const int nt1 = pad16(l) / 16, tiles = nt1 * nt1;
    for (int tile = wave; tile < 2 * tiles; tile += NW) {
      const int part = tile / tiles, i0 = ((tile % tiles) / nt1) * 16, j0 = (tile % nt1) * 16;
      const double* y = part == 0 ? sB : sG;
      d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
      for (int k0 = 0; k0 < kStepRows; k0 += 4) {
        const int k = k0 + lane / 16;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sB[k * ldr + i0 + lane % 16],
                                                   y[k * ldr + j0 + lane % 16], acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + lane / 16 + 4 * r, j = j0 + lane % 16;
        if (i < l && j < l) Pb[part * ll + i * l + j] = acc[r];
      }
    }
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) last = (atomicAdd(&st->ticket, 1) == static_cast<int>(nblk) - 1);
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (tid == 0) st->ticket = 0;  // ready for the next iteration's launch
  kn_step_ns<kStepThreads>(t, st, l, invN, tol, P, nblk, F, Cinv, smem, red);
}

// B+ = Bc F + Bg Cinv; E+ = Ec F + invN Bc Cinv; FF+ = FFc F (block 0).  Workgroup `blk` of
// nblocks (kThreads threads), a grid-stride loop over the elements: every element is one dot
// product in a fixed order, whatever the grid.
__device__ __forceinline__ void kn_update_body(
    int t, const IntegratorState* st, int64_t nl, int l, double invN,
    const double* __restrict__ Bc, const double* __restrict__ Bg, double* __restrict__ Bp,
    const double* __restrict__ Ec, double* __restrict__ Ep, const double* __restrict__ FFc,
    double* __restrict__ FFp, const double* __restrict__ F, const double* __restrict__ Cinv,
    int64_t blk, int64_t nblocks) {
  if (skip(st, t)) return;
  __shared__ double sF[kMaxL * kMaxL];
  __shared__ double sC[kMaxL * kMaxL];
  for (int e = threadIdx.x; e < l * l; e += kThreads) {
    sF[e] = F[e];
    sC[e] = Cinv[e];
  }
  __syncthreads();
  const int64_t total = nl * l;
  for (int64_t e = blk * kThreads + threadIdx.x; e < total; e += nblocks * kThreads) {
    const int64_t r = e / l;
    const int j = e % l;
    const double* bc = Bc + r * l;
    const double* bg = Bg + r * l;
    const double* ec = Ec + r * l;
    double b = 0.0, ee = 0.0;
    for (int q = 0; q < l; ++q) {
      const double f = sF[q * l + j], ci = sC[q * l + j];
      b += bc[q] * f + bg[q] * ci;
      ee += ec[q] * f + invN * bc[q] * ci;
    }
    Bp[e] = b;
    Ep[e] = ee;
  }
  if (blk == 0) {
    for (int e = threadIdx.x; e < l * l; e += kThreads) {
      const int i = e / l, j = e % l;
      double acc = 0.0;
      for (int q = 0; q < l; ++q) acc += FFc[i * l + q] * sF[q * l + j];
      FFp[e] = acc;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// KN iterations of several solves per launch.  A group's launch runs workgroup x of member y
// (blockIdx.y) on that member's own buffers, shapes, state and iteration index, with the
// workgroup -> rows/tiles map the member would have alone, so a member's arithmetic (and result,
// bit for bit) does not depend on the group it runs in or on when it joined.  A solve alone is a
// group of one.  The member table is a
// by-value kernel argument (read through the kernel-argument segment), so a launch needs no
// device memory of its own and can be recorded into a HIP graph as is.

struct KnMember {
  const double* Bs;   // Nl x Nl
  double* Pbg;        // bs_splits split-K partials of Bs * B
  double* B[2];
  double* Bg;
  double* E[2];
  double* FF[2];
  double* P;          // block partials of the step
  double* F;
  double* C;          // Cinv
  IntegratorState* st;
  int64_t nl, kchunk, bs_splits, bs_blocks, nblk_s, upd_blocks;
  double invN, tol;
  int l, vec;
  int t;              // the member's iteration index in this launch
};

struct KnGroup {
  int count;
  KnMember m[kMaxKnGroup];
};
static_assert(sizeof(KnGroup) <= 4096, "the member table fits the kernel arguments");

// Bg partials of iteration t: Pbg[s] = Bs(:, K range s) * Bc(K range s, :) on the tall-skinny
// FP64-MFMA tiles (the alpha = 1/N scaling is applied where the partials are summed).
template <int LP>
__global__ __launch_bounds__(256, LP == 32 ? 3 : 2) void kn_bs_group_kernel(KnGroup g) {
  const KnMember& m = g.m[blockIdx.y];
  const int t = m.t;
  if (blockIdx.x >= m.bs_blocks || skip(m.st, t)) return;
  GemmF64Args p;
  p.M = m.nl;
  p.N = m.l;
  p.K = m.nl;
  p.A = m.Bs;
  p.lda = m.nl;
  p.a_mcontig = false;
  p.B = m.B[t & 1];
  p.ldb = m.l;
  p.splits = m.bs_splits;
  p.kchunk = m.kchunk;
  p.partial = m.Pbg;
  if (m.vec)
    tall_direct_tile<double, double, LP, true>(p, blockIdx.x);
  else
    tall_direct_tile<double, double, LP, false>(p, blockIdx.x);
}

__global__ __launch_bounds__(kStepThreads) void kn_reduce_step_group_kernel(KnGroup g) {
  const KnMember& m = g.m[blockIdx.y];
  const int t = m.t;
  if (blockIdx.x >= m.nblk_s) return;
  kn_reduce_step_body(t, m.st, m.nl, m.l, m.invN, m.tol, m.Pbg, m.bs_splits, m.B[t & 1], m.Bg,
                      m.P, m.nblk_s, m.F, m.C, blockIdx.x);
}

__global__ __launch_bounds__(kThreads) void kn_update_group_kernel(KnGroup g) {
  const KnMember& m = g.m[blockIdx.y];
  const int t = m.t;
  if (blockIdx.x >= m.upd_blocks) return;
  const int c = t & 1, p = c ^ 1;
  kn_update_body(t, m.st, m.nl, m.l, m.invN, m.B[c], m.Bg, m.B[p], m.E[c], m.E[p], m.FF[c],
                 m.FF[p], m.F, m.C, blockIdx.x, m.upd_blocks);
}

// out[2i] = converged_at, out[2i + 1] = nan_seen of member i (one read-back for the group).
__global__ void kn_flags_kernel(KnGroup g, int* out) {
  const int i = threadIdx.x;
  if (i >= g.count) return;
  out[2 * i] = g.m[i].st->converged_at;
  out[2 * i + 1] = g.m[i].st->nan_seen;
}

// ---------------------------------------------------------------------------------------------
// Wen-Yin (row_block_gramian_wen_yin_integrator.hpp:162-317).  The line search needs
// ||Bc Fc + Bg Fgc||_F^2 per trial step; with Dc = Bc'Bc/N, Dgc = Bc'Bg/N and Hgg = Bg'Bg/N it is
// N * (tr(Fc'Dc Fc) + 2 tr(Fc'Dgc Fgc) + tr(Fgc'Hgg Fgc)), so every trial is l x l work.

struct WyConst {
  static constexpr double tau0 = 1.0, taumax = 1e14, taumin = 1e-14, beta = 0.5, sigma = 1e-4,
                          eta = 0.85;
  static constexpr int taumaxiter = 20;
};

// The l x l Wen-Yin kernels run on one workgroup of 16 waves.  An l x l matrix (l <= 64, padded
// with zeros to np = pad16(l)) is cut into (np/16)^2 tiles of 16 x 16: wave w owns tile
// (w / nt1, w % nt1), nt1 = np/16, and lane ln its entries (i0 + ln/16 + 4r, j0 + ln%16),
// r = 0..3 -- the layout of a v_mfma_f64_16x16x4f64 result, so a product lands in the registers
// of the entries it defines.  Four np x (np + 1) LDS buffers.
constexpr int kWyWaves = 16;
constexpr int kWyThreads = kWyWaves * kWave;

size_t wy_lds_bytes(int l) {
  const size_t np = pad16(l);
  return 4 * np * (np + 1) * sizeof(double) + 2 * kMaxL * sizeof(int) + kWyWaves * sizeof(double);
}

struct WyTile {
  int i0, j0, oj, ib;  // tile origin, owned column, first owned row
  bool act;            // the wave owns a tile
  __device__ WyTile(int np) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave, nt1 = np / 16;
    act = wave < nt1 * nt1;
    i0 = act ? (wave / nt1) * 16 : 0;
    j0 = act ? (wave % nt1) * 16 : 0;
    oj = j0 + (lane & 15);
    ib = i0 + (lane >> 4);
  }
  __device__ int row(int r) const { return ib + 4 * r; }
};

// Gauss-Jordan inverse with partial pivoting of the l x l matrix in X (LDS, leading dimension
// ld): the inverse LAPACK getrf+getri computes (getrfi.hpp:117-124).  Rows are never swapped:
// step k eliminates column k with the unused row p_k of largest |x_ik| (first on ties) and keeps
// the new column of the inverse in place of column k, so on return M[p_k][q] with p_q = j holds
// inv[k][j] (piv[k] = p_k) -- the same operations as with explicit row and column interchanges.
// Each step reads one buffer and writes the other, so it needs one barrier; wave 0 also forms
// the next pivot column and picks the next pivot.  Every thread must call this; returns the
// buffer that holds M.
__device__ double* wy_gj_inverse(int l, int ld, double* X, double* Y, int* piv, const WyTile& w) {
  const int lane = threadIdx.x % kWave;
  const bool wave0 = threadIdx.x < kWave;
  bool used = false;  // wave 0: row `lane` was a pivot
  // pick the pivot of column k among the unused rows; y = |column k| of this lane's row
  auto choose = [&](double y, int k) {
    const bool cand = lane < l && !used;
    double key = cand ? fabs(y) : -1.0;
    if (cand && key != key) key = -0.5;  // NaN only when no finite candidate is left
    int idx = cand ? lane : kWave;
    for (int off = 32; off > 0; off >>= 1) {
      const double ok = __shfl_xor(key, off, kWave);
      const int oi = __shfl_xor(idx, off, kWave);
      if (ok > key || (ok == key && oi < idx)) {
        key = ok;
        idx = oi;
      }
    }
    if (lane == idx) used = true;
    if (lane == 0) piv[k] = idx;
  };
  if (wave0) choose(lane < l ? X[lane * ld] : 0.0, 0);
  __syncthreads();
  for (int k = 0; k < l; ++k) {
    const int p = piv[k];
    const double pv = 1.0 / X[p * ld + k];
    if (w.act && w.oj < l) {
      const int j = w.oj;
      const double rk = (j == k) ? pv : X[p * ld + j] * pv;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = w.row(r);
        if (i >= l) continue;
        if (i == p) {
          Y[i * ld + j] = rk;
        } else {
          const double f = X[i * ld + k];
          Y[i * ld + j] = (j == k) ? -f * rk : X[i * ld + j] - f * rk;
        }
      }
    }
    if (wave0 && k + 1 < l) {
      double y = 0.0;
      if (lane < l) {
        const double rk1 = X[p * ld + k + 1] * pv;
        y = (lane == p) ? rk1 : X[lane * ld + k + 1] - X[lane * ld + k] * rk1;
      }
      choose(y, k + 1);
    }
    __syncthreads();
    double* z = X;
    X = Y;
    Y = z;
  }
  return X;
}

// Line search + phi/zeta update (row_block_gramian_wen_yin_integrator.hpp:181-222).  D3 = [Dc;
// Dgc; Hgg] of parity c.  Writes Fc, Fgc.
//
// The reference inverts C = [[Dc/2 - I/tau, I/2], [-Dgc/2, -Dc/2 - I/tau]] (2l x 2l) for every
// trial tau and takes Fgc = X12 Dc - X11, Fc = X22 Dc - X21 + I from X = C^{-1}, i.e.
// [Fgc; Fc - I] = C^{-1} [-I; Dc].  Eliminating the second block row of that solve gives
//   Fgc = (4/tau) (K + (4/tau^2) I)^{-1},  Fc = -I - (Dc - s I) Fgc,  K = Dgc - Dc^2, s = 2/tau,
// one l x l inverse per trial.  Its test value (1/2N) ||Bc Fc + Bg Fgc||_F^2 with Dc = Bc'Bc/N,
// Dgc = Bc'Bg/N, Hgg = Bg'Bg/N is, with F = Fgc,
//   phit = 1/2 [tr(Dc) - 2 tr((K + s Dc) F) + tr(F' X(s) F)],  X(s) = X0 + s X1 + s^2 Dc,
//   X0 = Dc^3 - Dc Dgc - (Dc Dgc)' + Hgg,  X1 = Dgc + Dgc' - 2 Dc^2,
// so a trial costs one inverse and one product.  K, X0 and X1 stay in registers (owned entries).
__global__ __launch_bounds__(kWyThreads) void wy_search_kernel(int t, IntegratorState* st, int l,
                                                              const double* __restrict__ D3,
                                                              double* __restrict__ Fc,
                                                              double* __restrict__ Fgc) {
  if (skip(st, t)) return;
  extern __shared__ double dyn[];
  __shared__ double phit_s, trdc_s;
  __shared__ int accept_s;
  const int np = pad16(l), ld = np + 1, nb = np * ld;
  double* B0 = dyn;  // Dc
  double* B1 = B0 + nb;
  double* B2 = B1 + nb;
  double* B3 = B2 + nb;
  int* piv = reinterpret_cast<int*>(B3 + nb);
  int* pinv = piv + kMaxL;
  double* red = reinterpret_cast<double*>(pinv + kMaxL);
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const WyTile w(np);
  const int oj = w.oj;
  const double* Dc = D3;
  const double* Dgc = D3 + l * l;
  const double* Hgg = D3 + 2 * l * l;
  for (int e = tid; e < np * np; e += kWyThreads) {
    const int i = e / np, j = e % np;
    const bool in = i < l && j < l;
    B0[i * ld + j] = in ? Dc[i * l + j] : 0.0;
    B1[i * ld + j] = in ? Dgc[i * l + j] : 0.0;
  }
  __syncthreads();
  if (w.act) {  // Dc^2 into B2
    const d4 a = lds_mfma_tile<false, false>(np, ld, B0, B0, w.i0, w.j0);
#pragma unroll
    for (int r = 0; r < 4; ++r) B2[w.row(r) * ld + oj] = a[r];
  }
  __syncthreads();
  double K[4] = {0, 0, 0, 0}, X0[4] = {0, 0, 0, 0}, X1[4] = {0, 0, 0, 0};
  if (w.act) {
    const d4 c3 = lds_mfma_tile<false, false>(np, ld, B0, B2, w.i0, w.j0);  // Dc^3
    const d4 cg = lds_mfma_tile<false, false>(np, ld, B0, B1, w.i0, w.j0);  // Dc Dgc
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = w.row(r);
      const double g = B1[i * ld + oj], d2 = B2[i * ld + oj];
      K[r] = g - d2;
      X1[r] = g + B1[oj * ld + i] - 2.0 * d2;
      X0[r] = c3[r];
      B3[i * ld + oj] = cg[r];
    }
  }
  if (wave == 0) {
    const double d = wave_sum(lane < l ? B0[lane * ld + lane] : 0.0);
    if (lane == 0) trdc_s = d;
  }
  __syncthreads();
  if (w.act) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = w.row(r);
      const double h = (i < l && oj < l) ? Hgg[i * l + oj] : 0.0;
      X0[r] = X0[r] - B3[i * ld + oj] - B3[oj * ld + i] + h;
    }
  }
  __syncthreads();  // B1..B3 are free from here
  const double phi = st->phi, mu = st->mu, trdc = trdc_s;
  double tau = st->taug, phit = phi, s = 2.0 / tau;
  double* Fb = B2;
  for (int ti = 0; ti < WyConst::taumaxiter; ++ti, tau *= WyConst::beta) {
    const double c = 4.0 / (tau * tau);
    s = 2.0 / tau;
    if (w.act) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = w.row(r);
        B1[i * ld + oj] = K[r] + (i == oj ? c : 0.0);
      }
    }
    __syncthreads();
    const double* R = wy_gj_inverse(l, ld, B1, B2, piv, w);
    Fb = (R == B1) ? B2 : B1;
    if (tid < l) pinv[piv[tid]] = tid;
    __syncthreads();
    if (w.act) {  // F = (4/tau) a^{-1} (gathered from the pivoted form), Xs = X0 + s X1 + s^2 Dc
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = w.row(r);
        Fb[i * ld + oj] = (i < l && oj < l) ? (4.0 / tau) * R[piv[i] * ld + pinv[oj]] : 0.0;
        B3[i * ld + oj] = X0[r] + s * X1[r] + s * s * B0[i * ld + oj];
      }
    }
    __syncthreads();
    double part = 0.0;
    if (w.act) {
      const d4 xf = lds_mfma_tile<false, false>(np, ld, B3, Fb, w.i0, w.j0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = w.row(r);
        part += Fb[i * ld + oj] * xf[r] - 2.0 * (K[r] + s * B0[i * ld + oj]) * Fb[oj * ld + i];
      }
    }
    part = wave_sum(part);
    if (lane == 0) red[wave] = part;
    __syncthreads();
    if (tid == 0) {
      double sum = trdc;
      for (int q = 0; q < kWyWaves; ++q) sum += red[q];
      phit_s = 0.5 * sum;
      accept_s = (phit_s >= phi + tau * WyConst::sigma * mu) ? 1 : 0;
    }
    __syncthreads();
    phit = phit_s;
    if (accept_s) break;
  }
  if (w.act) {  // Fgc = F, Fc = s F - Dc F - I
    const d4 df = lds_mfma_tile<false, false>(np, ld, B0, Fb, w.i0, w.j0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = w.row(r);
      if (i < l && oj < l) {
        const double f = Fb[i * ld + oj];
        Fgc[i * l + oj] = f;
        Fc[i * l + oj] = s * f - df[r] - (i == oj ? 1.0 : 0.0);
      }
    }
  }
  if (tid == 0) {
    const double zeta = st->zeta;
    st->phi = (WyConst::eta * zeta * phi + phit) / (WyConst::eta * zeta + 1.0);
    st->zeta = WyConst::eta * zeta + 1.0;
  }
}

// B+ = Bc Fc + Bg Fgc; U1 = Bc Fgc / N; BsU1 = Bg Fgc (= Bs U1); E+ = U1 + Ec Fc; FF+ = FFc Fc.
__global__ __launch_bounds__(kThreads) void wy_update_kernel(
    int t, const IntegratorState* st, int64_t nl, int l, double invN,
    const double* __restrict__ Bc, const double* __restrict__ Bg, double* __restrict__ Bp,
    double* __restrict__ U1, double* __restrict__ BsU1, const double* __restrict__ Ec,
    double* __restrict__ Ep, const double* __restrict__ FFc, double* __restrict__ FFp,
    const double* __restrict__ Fc, const double* __restrict__ Fgc) {
  if (skip(st, t)) return;
  // Workgroup b updates rows [kURows b, kURows b + kURows): their rows of Bc, Bg, Ec are staged
  // in LDS (coalesced), lane j of wave w computes column j of rows w, w + 4, ... (F and G are
  // read once per k for all of them); each sum runs over k in order.  The first
  // ceil(l^2 / kThreads) workgroups also form FF+ = FFc F, one entry per thread.
  __shared__ double sF[kMaxL * kMaxL];
  __shared__ double sG[kMaxL * kMaxL];
  __shared__ double sR[3][kURows][kMaxL];
  constexpr int RW = kURows / (kThreads / kWave);  // rows per wave
  const int tid = threadIdx.x;
  for (int e = tid; e < l * l; e += kThreads) {
    sF[e] = Fc[e];
    sG[e] = Fgc[e];
  }
  const int64_t r0 = blockIdx.x * static_cast<int64_t>(kURows);
  for (int e = tid; e < kURows * l; e += kThreads) {
    const int rr = e / l, j = e % l;
    const int64_t r = r0 + rr;
    const bool ok = r < nl;
    sR[0][rr][j] = ok ? Bc[r * l + j] : 0.0;
    sR[1][rr][j] = ok ? Bg[r * l + j] : 0.0;
    sR[2][rr][j] = ok ? Ec[r * l + j] : 0.0;
  }
  __syncthreads();
  const int j = tid % kWave, w = tid / kWave;
  if (j < l && r0 < nl) {
    double bcf[RW], bgg[RW], bcg[RW], ecf[RW];
#pragma unroll
    for (int u = 0; u < RW; ++u) bcf[u] = bgg[u] = bcg[u] = ecf[u] = 0.0;
    for (int q = 0; q < l; ++q) {
      const double f = sF[q * l + j], g = sG[q * l + j];
#pragma unroll
      for (int u = 0; u < RW; ++u) {
        const int rr = w + 4 * u;
        const double bc = sR[0][rr][q], bg = sR[1][rr][q], ec = sR[2][rr][q];
        bcf[u] += bc * f;
        bgg[u] += bg * g;
        bcg[u] += bc * g;
        ecf[u] += ec * f;
      }
    }
#pragma unroll
    for (int u = 0; u < RW; ++u) {
      const int64_t r = r0 + w + 4 * u;
      if (r < nl) {
        const int64_t e = r * l + j;
        Bp[e] = bcf[u] + bgg[u];
        U1[e] = invN * bcg[u];
        BsU1[e] = bgg[u];
        Ep[e] = invN * bcg[u] + ecf[u];
      }
    }
  }
  // FF+: this workgroup's entries [kThreads b, kThreads (b + 1)) lie in rows
  // [kThreads b / l, ...] of FFc: at most kThreads / l + 1 rows, staged like the others
  const int64_t fe0 = blockIdx.x * static_cast<int64_t>(kThreads);
  if (fe0 < static_cast<int64_t>(l) * l) {
    __syncthreads();  // sR is reused
    const int i0 = static_cast<int>(fe0 / l);
    const int nr = std::min(l - i0, kThreads / l + 1);
    double* sFF = &sR[0][0][0];
    for (int e = tid; e < nr * l; e += kThreads) sFF[e] = FFc[i0 * l + e];
    __syncthreads();
    const int64_t fe = fe0 + tid;
    if (fe < static_cast<int64_t>(l) * l) {
      const int i = static_cast<int>(fe / l), jj = static_cast<int>(fe % l);
      double acc = 0.0;
      for (int q = 0; q < l; ++q) acc += sFF[(i - i0) * l + q] * sF[q * l + jj];
      FFp[fe] = acc;
    }
  }
}

// mu = sum |diag(Dg+)| - ||D+||_F^2; stop when mu < tol^2 (row_block_gramian_wen_yin_integrator.hpp:247-255).
__global__ void wy_check_kernel(int t, IntegratorState* st, int l, double tol,
                                const double* __restrict__ D3p) {
  if (skip(st, t)) return;
  __shared__ double red[kThreads / kWave];
  const double* Dp = D3p;
  const double* Dgp = D3p + l * l;
  double part = 0.0;
  for (int e = threadIdx.x; e < l * l; e += blockDim.x) {
    const int i = e / l, j = e % l;
    part -= Dp[e] * Dp[e];
    if (i == j) part += fabs(Dgp[e]);
  }
  part = wave_sum(part);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mu = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) mu += red[w];
    st->mu = mu;
    st->res = mu;
    if (!isfinite(mu)) st->nan_seen = 1;
    if (mu < tol * tol) st->converged_at = t;
  }
}

// Nl-row pass for the step-size terms (row_block_gramian_wen_yin_integrator.hpp:260-312):
// U2 = (B+ - Bc)/N - U1 D+, BsU = Bs U1 (odd) or Bs U2 = (Bg+ - Bgc) - BsU1 D+ (even);
// per-block partials [Bc'U1, Bc'U2, Bc'Mc (l x l each), U1.BsU, U2.BsU], where
// Mc = E~c + [F~c; 0] are the coefficients of Qc = Qs Mc, so Bc'Mc = Qc'Qc.
__global__ __launch_bounds__(kThreads) void wy_terms_kernel(
    int t, const IntegratorState* st, int64_t nl, int l, double invN, int odd,
    const double* __restrict__ Bc, const double* __restrict__ Bp, const double* __restrict__ Bgc,
    const double* __restrict__ Bgp, const double* __restrict__ U1,
    const double* __restrict__ BsU1, const double* __restrict__ Dp,
    const double* __restrict__ Ec, const double* __restrict__ FFc, double* __restrict__ P) {
  if (skip(st, t) || st->converged_at == t) return;
  __shared__ double sD[kMaxL * kMaxL];
  __shared__ double sBc[kRows][kMaxL + 1];  // rows padded: the MFMA fragment reads below
  __shared__ double sU1[kRows][kMaxL + 1];  // take two k-rows per 32 lanes
  __shared__ double sU2[kRows][kMaxL + 1];
  __shared__ double sBsU[kRows][kMaxL];
  __shared__ double sM[kRows][kMaxL + 1];
  __shared__ double sBsU1[kRows][kMaxL];
  __shared__ double red[2][kThreads / kWave];
  for (int e = threadIdx.x; e < l * l; e += kThreads) sD[e] = Dp[e];
  const int64_t r0 = blockIdx.x * (int64_t)kRows;
  // the block's rows of U1 and BsU1 go to LDS first (coalesced): the products with D+ below
  // read them along the row
  for (int e = threadIdx.x; e < kRows * l; e += kThreads) {
    const int rr = e / l, j = e % l;
    const int64_t r = r0 + rr;
    sU1[rr][j] = (r < nl) ? U1[r * l + j] : 0.0;
    sBsU1[rr][j] = (r < nl) ? BsU1[r * l + j] : 0.0;
  }
  // columns l .. pad16(l) of the MFMA operands are zero (those tiles' outputs are not stored)
  for (int e = threadIdx.x; e < kRows * (kMaxL - l); e += kThreads) {
    const int rr = e / (kMaxL - l), j = l + e % (kMaxL - l);
    sBc[rr][j] = 0.0;
    sU1[rr][j] = 0.0;
    sU2[rr][j] = 0.0;
    sM[rr][j] = 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kRows * l; e += kThreads) {
    const int rr = e / l, j = e % l;
    const int64_t r = r0 + rr;
    double bc = 0, u2 = 0, bsu = 0, mc = 0;
    if (r < nl) {
      const int64_t o = r * l;
      bc = Bc[o + j];
      mc = Ec[o + j] + (r < l ? FFc[o + j] : 0.0);
      double u1d = 0.0, bsu1d = 0.0;
      for (int q = 0; q < l; ++q) {
        u1d += sU1[rr][q] * sD[q * l + j];
        bsu1d += sBsU1[rr][q] * sD[q * l + j];
      }
      u2 = invN * (Bp[o + j] - bc) - u1d;
      bsu = odd ? sBsU1[rr][j] : (Bgp[o + j] - Bgc[o + j]) - bsu1d;
    }
    sBc[rr][j] = bc;
    sU2[rr][j] = u2;
    sBsU[rr][j] = bsu;
    sM[rr][j] = mc;
  }
  __syncthreads();
  const int count = 3 * l * l + 2;
  double* Pb = P + blockIdx.x * (int64_t)count;
  {  // Bc'U1, Bc'U2, Bc'Mc over the block's rows on f64 MFMA tiles (16 x 16, k = row)
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int nt1 = pad16(l) / 16, tiles = nt1 * nt1;
    for (int tile = wave; tile < 3 * tiles; tile += kThreads / kWave) {
      const int part = tile / tiles, i0 = ((tile % tiles) / nt1) * 16, j0 = (tile % nt1) * 16;
      const double(*Y)[kMaxL + 1] = part == 0 ? sU1 : (part == 1 ? sU2 : sM);
      d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
      for (int k0 = 0; k0 < kRows; k0 += 4) {
        const int k = k0 + lane / 16;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sBc[k][i0 + lane % 16], Y[k][j0 + lane % 16],
                                                   acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + lane / 16 + 4 * r, j = j0 + lane % 16;
        if (i < l && j < l) Pb[part * l * l + i * l + j] = acc[r];
      }
    }
  }
  double d1 = 0.0, d2 = 0.0;
  for (int e = threadIdx.x; e < kRows * l; e += kThreads) {
    const int rr = e / l, j = e % l;
    d1 += sU1[rr][j] * sBsU[rr][j];
    d2 += sU2[rr][j] * sBsU[rr][j];
  }
  d1 = wave_sum(d1);
  d2 = wave_sum(d2);
  if (threadIdx.x % kWave == 0) {
    red[0][threadIdx.x / kWave] = d1;
    red[1][threadIdx.x / kWave] = d2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s1 = 0, s2 = 0;
    for (int w = 0; w < kThreads / kWave; ++w) {
      s1 += red[0][w];
      s2 += red[1][w];
    }
    Pb[3 * l * l] = s1;
    Pb[3 * l * l + 1] = s2;
  }
}

// t1, t2 and the Barzilai-Borwein step (row_block_gramian_wen_yin_integrator.hpp:260-316), with
// T = [Bc'U1; Bc'U2; Qc'Qc; U1.BsU; U2.BsU] (reduced, unscaled).  The quadratic terms in Phi use
// the actual Qc'Qc where the reference's Gramian form assumes the identity: ||dQ||^2 =
// ||Qc Phi1 + Qs U1||^2 then equals the row-block form's explicit sum of squares, and the step
// size no longer amplifies the drift of Qc'Qc from I (with the identity, a 1e-15 perturbation of
// the input moves the result by ~1e-5 at the CPU-runnable workload).
__global__ __launch_bounds__(kWyThreads) void wy_tau_kernel(int t, IntegratorState* st, int l,
                                                           int odd, const double* __restrict__ Dc,
                                                           const double* __restrict__ Dp,
                                                           const double* __restrict__ Fc,
                                                           const double* __restrict__ T) {
  if (skip(st, t) || st->converged_at == t) return;
  extern __shared__ double dyn[];
  __shared__ double red7[7][kWyWaves];
  const int np = pad16(l), ld = np + 1, nb = np * ld;
  double* sF = dyn;        // Fc
  double* sD = sF + nb;    // D+
  double* sQ = sD + nb;    // Qc'Qc
  double* sP2 = sQ + nb;   // Phi2 = Dc - Fc D+
  const double* BtU1 = T;
  const double* BtU2 = T + l * l;
  const double* QQ = T + 2 * l * l;
  const double u1bsu = T[3 * l * l], u2bsu = T[3 * l * l + 1];
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const WyTile w(np);
  const int oj = w.oj;
  for (int e = tid; e < np * np; e += kWyThreads) {
    const int i = e / np, j = e % np;
    const bool in = i < l && j < l;
    sF[i * ld + j] = in ? Fc[i * l + j] : 0.0;
    sD[i * ld + j] = in ? Dp[i * l + j] : 0.0;
    sQ[i * ld + j] = in ? QQ[i * l + j] : 0.0;
  }
  __syncthreads();
  if (w.act) {
    const d4 fd = lds_mfma_tile<false, false>(np, ld, sF, sD, w.i0, w.j0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = w.row(r);
      sP2[i * ld + oj] = ((i < l && oj < l) ? Dc[i * l + oj] : 0.0) - fd[r];
    }
  }
  __syncthreads();
  // <X, QQ Y> = sum_ij X_ij (QQ Y)_ij with Phi1 = Fc - I, so QQ Phi1 = QQ Fc - QQ
  double v[7] = {0, 0, 0, 0, 0, 0, 0};  // p11, p1bu1, p12, p1bu2, p2bu1, p22, p2bu2
  if (w.act) {
    const d4 q1 = lds_mfma_tile<false, false>(np, ld, sQ, sF, w.i0, w.j0);
    const d4 q2 = lds_mfma_tile<false, false>(np, ld, sQ, sP2, w.i0, w.j0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = w.row(r);
      if (i < l && oj < l) {
        const double ph1 = sF[i * ld + oj] - (i == oj ? 1.0 : 0.0), ph2 = sP2[i * ld + oj];
        const double qp1 = q1[r] - sQ[i * ld + oj];
        const double bu1 = BtU1[i * l + oj], bu2 = BtU2[i * l + oj];
        v[0] += ph1 * qp1;
        v[1] += ph1 * bu1;
        v[2] += ph1 * q2[r];
        v[3] += ph1 * bu2;
        v[4] += ph2 * bu1;
        v[5] += ph2 * q2[r];
        v[6] += ph2 * bu2;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const double x = wave_sum(v[k]);
    if (lane == 0) red7[k][wave] = x;
  }
  __syncthreads();
  if (tid == 0) {
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 7; ++k)
      for (int q = 0; q < kWyWaves; ++q) s[k] += red7[k][q];
    double t1, t2;
    if (odd) {
      t1 = s[0] + 2.0 * s[1] + u1bsu;
      t2 = s[2] + s[3] + s[4] + u2bsu;
    } else {
      t1 = s[2] + s[3] + s[4] + u1bsu;
      t2 = s[5] + 2.0 * s[6] + u2bsu;
    }
    double tg = fabs(t1 / t2);
    if (tg < WyConst::taumin) tg = WyConst::taumin;
    if (tg > WyConst::taumax) tg = WyConst::taumax;
    st->taug = tg;
  }
}

// ---------------------------------------------------------------------------------------------
__global__ void init_state_kernel(IntegratorState* st) {
  st->converged_at = INT_MAX;
  st->nan_seen = 0;
  st->ticket = 0;
  st->res = 0.0;
  st->taug = WyConst::tau0;
  st->zeta = 1.0;
  st->phi = 0.0;
  st->mu = 0.0;
}

// B0 = first l columns of Bs; E0 = 0; FF0 = I.
__global__ void init_coef_kernel(const double* __restrict__ Bs, int64_t nl, int l,
                                 double* __restrict__ B0, double* __restrict__ E0,
                                 double* __restrict__ FF0) {
  const int64_t total = nl * l;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / l;
    const int j = e % l;
    B0[e] = Bs[r * nl + j];
    E0[e] = 0.0;
    if (e < (int64_t)l * l) FF0[e] = ((e / l) == (e % l)) ? 1.0 : 0.0;
  }
}

// WY initial scalars: phi = ||Bc||^2 / 2N = tr(Dc)/2, mu = sum|diag(Dgc)| - ||Dc||^2.
// One wave.
__global__ void wy_init_scalars_kernel(IntegratorState* st, int l, const double* __restrict__ D3) {
  const double* Dc = D3;
  const double* Dgc = D3 + l * l;
  const int lane = threadIdx.x;
  double tr = 0.0, dg = 0.0, sq = 0.0;
  for (int i = lane; i < l; i += kWave) {
    tr += Dc[i * l + i];
    dg += fabs(Dgc[i * l + i]);
  }
  for (int e = lane; e < l * l; e += kWave) sq += Dc[e] * Dc[e];
  tr = wave_sum(tr);
  dg = wave_sum(dg);
  sq = wave_sum(sq);
  if (lane == 0) {
    st->phi = 0.5 * tr;
    st->mu = dg - sq;
  }
}

// Final re-orthonormalization of the basis Qbar = Qs M: Gm = M' Bs M = N * sum_blk P[blk][1]
// (P from gram_partials_kernel over M and Bs M / N), W = Gm^{-1/2} by the Newton-Schulz square
// root; Qbar W is the orthonormal basis nearest to Qbar (Loewdin).  The iterations keep Qbar
// orthonormal only to the rounding they accumulate (as the reference's own updates do); this
// step removes it.  W = I when Gm is not positive definite.
// (16 waves: one 16 x 16 tile of each product per wave at l = 64.)
__global__ __launch_bounds__(kWyThreads) void polish_kernel(int l, double N,
                                                           const double* __restrict__ P,
                                                           int64_t nblk, double* __restrict__ Wout) {
  __shared__ double sY[kMaxL * (kMaxL + 1)];
  __shared__ double sW[kMaxL * (kMaxL + 1)];
  __shared__ double sT[kMaxL * (kMaxL + 1)];
  __shared__ double red[kWyWaves];
  const int np = pad16(l), ld = np + 1, ll = l * l;
  for (int e = threadIdx.x; e < np * np; e += kWyThreads) {
    const int i = e / np, j = e % np;
    sT[i * ld + j] = (i < l && j < l) ? N * sum_partials(P, nblk, 2 * ll, ll + i * l + j)
                                      : ((i == j) ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < np * np; e += kWyThreads) {
    const int i = e / np, j = e % np;
    sY[i * ld + j] = 0.5 * (sT[i * ld + j] + sT[j * ld + i]);
  }
  __syncthreads();
  const int ok = lds_ns_sqrt<kWyWaves>(np, ld, sY, sW, sT, red, 60);
  for (int e = threadIdx.x; e < ll; e += kWyThreads) {
    const int i = e / l, j = e % l;
    Wout[e] = ok > 0 ? sW[i * ld + j] : ((i == j) ? 1.0 : 0.0);
  }
}

// out = X W for X Nl x l and W l x l (row-major).
__global__ __launch_bounds__(kThreads) void right_mul_kernel(int64_t nl, int l,
                                                            const double* __restrict__ X,
                                                            const double* __restrict__ W,
                                                            double* __restrict__ out) {
  __shared__ double sW[kMaxL * kMaxL];
  for (int e = threadIdx.x; e < l * l; e += kThreads) sW[e] = W[e];
  __syncthreads();
  const int64_t total = nl * l;
  for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kThreads) {
    const int64_t r = e / l;
    const int j = e % l;
    const double* x = X + r * l;
    double acc = 0.0;
    for (int q = 0; q < l; ++q) acc += x[q] * sW[q * l + j];
    out[e] = acc;
  }
}

// M = E + [FF; 0]  (Nl x l).
__global__ void coef_kernel(const double* __restrict__ E, const double* __restrict__ FF,
                            int64_t nl, int l, double* __restrict__ M) {
  const int64_t total = nl * l;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    M[e] = E[e] + (e < (int64_t)l * l ? FF[e] : 0.0);
  }
}

// Workgroups of wy_update_kernel: the row blocks, and at least those FF+ = FFc F needs.
unsigned wy_update_blocks(int64_t nl, int l) {
  const int64_t b = std::max(ceil_div(nl, kURows), ceil_div(static_cast<int64_t>(l) * l, kThreads));
  return static_cast<unsigned>(b);
}

unsigned grid_elems(int64_t total) {
  int64_t b = ceil_div(total, kThreads);
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return static_cast<unsigned>(b);
}

struct Layout {
  int64_t nl;
  int l;
  int64_t nblk_p;  // blocks of gram_partials_kernel
  int64_t nblk_t;  // blocks of wy_terms_kernel
  int64_t splits;  // split-K factor of the Bs products
  double* B[2];
  double* Bg[2];
  double* E[2];
  double* FF[2];
  double* D3[2];   // [Dc; Dgc; Hgg] per parity
  double* F;       // KN F / WY Fc
  double* C;       // KN Cinv / WY Fgc
  double* U1;
  double* BsU1;
  double* P;       // partial dot products
  double* Pbg;     // split-K partials of the Bs products
  double* T;       // reduced WY terms, 3 l^2 + 2
  int* flags;      // a KN group's read-back of [converged_at, nan_seen] per member
  IntegratorState* st;
  size_t bytes;
};

Layout layout(char* base, int64_t N, int l) {
  Layout L{};
  L.nl = N * l;
  L.l = l;
  L.nblk_p = ceil_div(L.nl, kPRows);
  L.nblk_t = ceil_div(L.nl, kRows);
  {  // about 256 workgroups (512 for large Nl, where each split stays long), >= 128 k per split
    const int64_t tiles = ceil_div(L.nl, 128);
    int64_t sp = ceil_div(L.nl > 2048 ? 512 : 256, tiles);
    const int64_t cap = ceil_div(L.nl, 128);
    L.splits = sp < cap ? sp : cap;
    if (L.splits < 1) L.splits = 1;
  }
  const size_t nll = static_cast<size_t>(L.nl) * l, ll = static_cast<size_t>(l) * l;
  size_t off = 0;
  auto take = [&](size_t elems) {
    double* p = reinterpret_cast<double*>(base ? base + off : nullptr);
    off += ((elems * sizeof(double) + 255) / 256) * 256;
    return p;
  };
  for (int i = 0; i < 2; ++i) {
    L.B[i] = take(nll);
    L.Bg[i] = take(nll);
    L.E[i] = take(nll);
    L.FF[i] = take(ll);
    L.D3[i] = take(3 * ll);
  }
  L.F = take(ll);
  L.C = take(ll);
  L.U1 = take(nll);
  L.BsU1 = take(nll);
  const size_t pp = static_cast<size_t>(L.nblk_p) * 3 * ll;
  const size_t pt = static_cast<size_t>(L.nblk_t) * (3 * ll + 2);
  L.P = take(pp > pt ? pp : pt);
  L.Pbg = take(static_cast<size_t>(L.splits + 1) * nll);
  L.T = take(3 * ll + 2);
  L.flags = reinterpret_cast<int*>(take(kMaxKnGroup));  // 2 kMaxKnGroup ints
  L.st = reinterpret_cast<IntegratorState*>(base ? base + off : nullptr);
  off += 256;
  L.bytes = off;
  return L;
}

}  // namespace

size_t integrator_workspace_bytes(int64_t N, int l) { return layout(nullptr, N, l).bytes; }

IterationGraphs::~IterationGraphs() {
  for (auto& e : execs) (void)hipGraphExecDestroy(e.second);
  if (cap) (void)hipStreamDestroy(cap);
}

namespace {

// Bout = Bs * Bin / N (Bout null: only the split-K partials into L.Pbg), gated on t, through the
// GEMM dispatch (the Wen-Yin iterations and the final polish).
void bs_product(const GramIntegrator& g, const Layout& L, const double* Bin, double* Bout, int t,
                hipStream_t st) {
  const int64_t nl = L.nl;
  GemmF64Args p;
  p.M = nl;
  p.N = g.l;
  p.K = nl;
  p.A = g.Bs;
  p.lda = nl;
  p.a_mcontig = false;
  p.B = Bin;
  p.ldb = g.l;
  p.C = Bout;
  p.ldc = g.l;
  p.alpha = 1.0 / static_cast<double>(g.N);
  p.shape = g.l <= 32 ? GemmShape::kTallN32 : GemmShape::kTallN64;
  p.splits = L.splits;
  p.kchunk = ceil_div(ceil_div(nl, L.splits), 16) * 16;
  p.partial = L.Pbg;
  p.gate = &L.st->converged_at;
  p.gate_t = t;
  if (g.l > 32 && nl % 2 == 0) {
    // 256 x 64 DMA tiles of Bs read as M-contiguous (Bs is stored exactly symmetric, so its
    // rows stand in for its columns): 157 us for the 8192 x 8192 Bs at l = 64, against 175 us
    // streamed and 196 us register-staged
    p.a_mcontig = true;
    p.b_ncontig = true;
    p.shape = GemmShape::kSquare;
  } else {
    // Bs streamed into LDS by DMA (tall_stream_kernel; the register-staged kernel where its
    // operand conditions fail)
    p.kernel = 1;
  }
  gemm_f64(p, st);
}

void check_integrator(const GramIntegrator& g, const Layout& L) {
  MCNLA_REQUIRE(g.l >= 1 && g.l <= kMaxL, MCNLA_ERR_INVALID, "integrator needs 1 <= l <= 64");
  MCNLA_REQUIRE(g.N >= 1, MCNLA_ERR_INVALID, "integrator needs N >= 1");
  MCNLA_REQUIRE(g.ws_bytes >= L.bytes, MCNLA_ERR_INVALID, "integrator workspace too small");
}

// State reset, B0 = Bs(:, 0:l), E0 = 0, FF0 = I.
void init_iteration(const GramIntegrator& g, const Layout& L, hipStream_t st) {
  hipLaunchKernelGGL(init_state_kernel, dim3(1), dim3(1), 0, st, L.st);
  MCNLA_LAUNCHED();
  hipLaunchKernelGGL(init_coef_kernel, dim3(grid_elems(L.nl * g.l)), dim3(kThreads), 0, st, g.Bs,
                     L.nl, g.l, L.B[0], L.E[0], L.FF[0]);
  MCNLA_LAUNCHED();
}

// The coefficients after `iters` iterations, re-orthonormalized: coef = M W with
// M = E + [FF; 0] and W = (M' Bs M)^{-1/2}.
void finish_iteration(const GramIntegrator& g, const Layout& L, int64_t iters, double* coef,
                      hipStream_t st) {
  const int64_t nl = L.nl;
  const int l = g.l;
  const int fin = static_cast<int>(iters & 1);
  hipLaunchKernelGGL(coef_kernel, dim3(grid_elems(nl * l)), dim3(kThreads), 0, st, L.E[fin],
                     L.FF[fin], nl, l, L.U1);
  MCNLA_LAUNCHED();
  bs_product(g, L, L.U1, L.BsU1, -1, st);
  hipLaunchKernelGGL(gram_partials_kernel, dim3(static_cast<unsigned>(L.nblk_p)), dim3(kThreads),
                     0, st, -1, L.st, nl, l, L.U1, L.BsU1, L.P, 2);
  MCNLA_LAUNCHED();
  hipLaunchKernelGGL(polish_kernel, dim3(1), dim3(kWyThreads), 0, st, l, static_cast<double>(g.N),
                     L.P, L.nblk_p, L.F);
  MCNLA_LAUNCHED();
  hipLaunchKernelGGL(right_mul_kernel, dim3(grid_elems(nl * l)), dim3(kThreads), 0, st, nl, l, L.U1,
                     L.F, coef);
  MCNLA_LAUNCHED();
}

KnMember kn_member(const GramIntegrator& g, const Layout& L) {
  KnMember m{};
  const int64_t nl = L.nl;
  m.Bs = g.Bs;
  m.Pbg = L.Pbg;
  m.B[0] = L.B[0];
  m.B[1] = L.B[1];
  m.Bg = L.Bg[0];
  m.E[0] = L.E[0];
  m.E[1] = L.E[1];
  m.FF[0] = L.FF[0];
  m.FF[1] = L.FF[1];
  m.P = L.P;
  m.F = L.F;
  m.C = L.C;
  m.st = L.st;
  m.nl = nl;
  m.l = g.l;
  m.kchunk = ceil_div(ceil_div(nl, L.splits), 16) * 16;  // whole 16-deep K tiles per split
  m.bs_splits = ceil_div(nl, m.kchunk);
  m.bs_blocks = ceil_div(nl, 128) * m.bs_splits;
  m.nblk_s = ceil_div(nl, kStepRows);
  m.upd_blocks = grid_elems(nl * g.l);
  m.invN = 1.0 / static_cast<double>(g.N);
  m.tol = g.tol;
  m.vec = (reinterpret_cast<uintptr_t>(g.Bs) % 16 == 0 && nl % 2 == 0) ? 1 : 0;
  return m;
}

// One KN iteration of the members in `act` (member i at its iteration act.m[i].t): the split-K
// Bs products (a launch per tile-width class present), then the reductions + l x l steps, then
// the updates.
void launch_kn_iteration(const KnGroup& act, hipStream_t st) {
  KnGroup narrow, wide;
  narrow.count = wide.count = 0;
  int64_t bn = 0, bw = 0, bs = 0, bu = 0;
  for (int i = 0; i < act.count; ++i) {
    const KnMember& m = act.m[i];
    if (m.l <= 32) {
      narrow.m[narrow.count++] = m;
      bn = std::max(bn, m.bs_blocks);
    } else {
      wide.m[wide.count++] = m;
      bw = std::max(bw, m.bs_blocks);
    }
    bs = std::max(bs, m.nblk_s);
    bu = std::max(bu, m.upd_blocks);
  }
  MCNLA_REQUIRE(bn < (1LL << 31) && bw < (1LL << 31) && bs < (1LL << 31), MCNLA_ERR_INVALID,
                "integrator grid too large");
  if (narrow.count) {
    hipLaunchKernelGGL(kn_bs_group_kernel<32>, dim3(static_cast<unsigned>(bn), narrow.count),
                       dim3(256), 0, st, narrow);
    MCNLA_LAUNCHED();
  }
  if (wide.count) {
    hipLaunchKernelGGL(kn_bs_group_kernel<64>, dim3(static_cast<unsigned>(bw), wide.count),
                       dim3(256), 0, st, wide);
    MCNLA_LAUNCHED();
  }
  hipLaunchKernelGGL(kn_reduce_step_group_kernel, dim3(static_cast<unsigned>(bs), act.count),
                     dim3(kStepThreads), 0, st, act);
  MCNLA_LAUNCHED();
  hipLaunchKernelGGL(kn_update_group_kernel, dim3(static_cast<unsigned>(bu), act.count),
                     dim3(kThreads), 0, st, act);
  MCNLA_LAUNCHED();
}

// The KN loop of a group.  `graphs` (a group of one only) records each chunk once and replays it.
void run_kn(const GramIntegrator* gs, double* const* coefs, int count, hipStream_t st,
            IterationGraphs* graphs, int64_t* iterations, bool* converged, bool* numeric_ok) {
  MCNLA_REQUIRE(count >= 1 && count <= kMaxKnGroup, MCNLA_ERR_INVALID,
                "a KN group holds 1 to 16 solves");
  MCNLA_REQUIRE(graphs == nullptr || count == 1, MCNLA_ERR_INVALID,
                "iteration graphs are recorded for a single solve only");
  std::vector<Layout> Ls(count);
  KnGroup all;
  all.count = count;
  for (int i = 0; i < count; ++i) {
    MCNLA_REQUIRE(!gs[i].wen_yin, MCNLA_ERR_INVALID, "a KN group holds KN integrations only");
    Ls[i] = layout(gs[i].ws, gs[i].N, gs[i].l);
    check_integrator(gs[i], Ls[i]);
    all.m[i] = kn_member(gs[i], Ls[i]);
  }
  for (int i = 0; i < count; ++i) init_iteration(gs[i], Ls[i], st);

  std::vector<int> conv_at(count, INT_MAX), nan_seen(count, 0), flags(2 * count);
  int64_t t = 0;
  int64_t chunk = 8;
  hipStream_t ls = st;  // where the iteration launches go: a capture stream while recording
  for (;;) {
    int64_t tmax = 0;  // the furthest iteration any unconverged member may still run
    for (int i = 0; i < count; ++i)
      if (conv_at[i] == INT_MAX) tmax = std::max(tmax, gs[i].maxiter);
    if (t >= tmax) break;
    const int64_t tend = (t + chunk < tmax) ? t + chunk : tmax;
    hipGraphExec_t exec = graphs ? graphs->find(t, tend) : nullptr;
    const bool recording = exec == nullptr && graphs != nullptr;
    if (recording) {
      if (!graphs->cap) MCNLA_HIP(hipStreamCreateWithFlags(&graphs->cap, hipStreamNonBlocking));
      ls = graphs->cap;
      MCNLA_HIP(hipStreamBeginCapture(ls, hipStreamCaptureModeThreadLocal));
    }
    const int64_t t0 = t;
    if (exec != nullptr) {
      t = tend;
    } else try {
      for (; t < tend; ++t) {
        KnGroup act;  // members not known to have converged and below their maxiter
        act.count = 0;
        for (int i = 0; i < count; ++i)
          if (conv_at[i] == INT_MAX && t < gs[i].maxiter) {
            act.m[act.count] = all.m[i];
            act.m[act.count++].t = static_cast<int>(t);
          }
        launch_kn_iteration(act, ls);
      }
    } catch (...) {
      if (recording) {  // leave the capture stream usable
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(ls, &partial);
        if (partial) (void)hipGraphDestroy(partial);
      }
      throw;
    }
    if (recording) {
      hipGraph_t graph = nullptr;
      MCNLA_HIP(hipStreamEndCapture(ls, &graph));
      ls = st;
      hipGraphExec_t made = nullptr;
      const hipError_t e = hipGraphInstantiate(&made, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      MCNLA_HIP(e);
      graphs->execs.push_back({{t0, tend}, made});
      exec = made;
    }
    if (exec != nullptr) MCNLA_HIP(hipGraphLaunch(exec, st));
    hipLaunchKernelGGL(kn_flags_kernel, dim3(1), dim3(kWave), 0, st, all, Ls[0].flags);
    MCNLA_LAUNCHED();
    MCNLA_HIP(hipMemcpyAsync(flags.data(), Ls[0].flags, 2 * count * sizeof(int),
                             hipMemcpyDeviceToHost, st));
    MCNLA_HIP(hipStreamSynchronize(st));
    bool running = false;
    for (int i = 0; i < count; ++i) {
      conv_at[i] = flags[2 * i];
      nan_seen[i] = flags[2 * i + 1];
      if (conv_at[i] == INT_MAX && t < gs[i].maxiter) running = true;
    }
    if (!running) break;
    if (chunk < 32) chunk *= 2;
  }
  for (int i = 0; i < count; ++i) {
    const int64_t iters =
        (conv_at[i] != INT_MAX) ? static_cast<int64_t>(conv_at[i]) + 1 : gs[i].maxiter;
    finish_iteration(gs[i], Ls[i], iters, coefs[i], st);
    if (iterations) iterations[i] = iters;
    if (converged) converged[i] = conv_at[i] != INT_MAX;
    if (numeric_ok) numeric_ok[i] = nan_seen[i] == 0;
  }
}

}  // namespace

void integrate_gram_group(const GramIntegrator* gs, double* const* coefs, int count,
                          hipStream_t st, int64_t* iterations, bool* converged, bool* numeric_ok) {
  run_kn(gs, coefs, count, st, nullptr, iterations, converged, numeric_ok);
}

struct KnBatch::Impl {
  struct Slot {
    bool used = false;
    GramIntegrator g;
    Layout L;
    KnMember m;
    int64_t t = 0;
    int conv_at = INT_MAX;
    int nan_seen = 0;
  };
  int* flags = nullptr;
  Slot slot[kMaxKnGroup];
  int host[2 * kMaxKnGroup];
  bool live(const Slot& x) const {
    return x.used && x.conv_at == INT_MAX && x.t < x.g.maxiter;
  }
};

KnBatch::KnBatch(int* flags) : impl_(new Impl) { impl_->flags = flags; }
KnBatch::~KnBatch() { delete impl_; }

int KnBatch::free_slots() const {
  int n = 0;
  for (const auto& x : impl_->slot) n += x.used ? 0 : 1;
  return n;
}

int KnBatch::add(const GramIntegrator& g, hipStream_t st) {
  MCNLA_REQUIRE(!g.wen_yin, MCNLA_ERR_INVALID, "a KN batch holds KN integrations only");
  for (int i = 0; i < kMaxKnGroup; ++i) {
    Impl::Slot& x = impl_->slot[i];
    if (x.used) continue;
    x.g = g;
    x.L = layout(g.ws, g.N, g.l);
    check_integrator(g, x.L);
    x.m = kn_member(g, x.L);
    x.t = 0;
    x.conv_at = INT_MAX;
    x.nan_seen = 0;
    init_iteration(g, x.L, st);
    x.used = true;
    return i;
  }
  throw Error{MCNLA_ERR_STATE, "KN batch is full"};
}

void KnBatch::advance(int iters, hipStream_t st) {
  for (int k = 0; k < iters; ++k) {
    KnGroup act;
    act.count = 0;
    for (auto& x : impl_->slot)
      if (impl_->live(x)) {
        act.m[act.count] = x.m;
        act.m[act.count++].t = static_cast<int>(x.t);
        ++x.t;
      }
    if (act.count == 0) break;
    launch_kn_iteration(act, st);
  }
  KnGroup used;
  used.count = 0;
  int who[kMaxKnGroup];
  for (int i = 0; i < kMaxKnGroup; ++i)
    if (impl_->slot[i].used) {
      who[used.count] = i;
      used.m[used.count++] = impl_->slot[i].m;
    }
  if (used.count == 0) return;
  hipLaunchKernelGGL(kn_flags_kernel, dim3(1), dim3(kWave), 0, st, used, impl_->flags);
  MCNLA_LAUNCHED();
  MCNLA_HIP(hipMemcpyAsync(impl_->host, impl_->flags, 2 * used.count * sizeof(int),
                           hipMemcpyDeviceToHost, st));
  MCNLA_HIP(hipStreamSynchronize(st));
  for (int j = 0; j < used.count; ++j) {
    Impl::Slot& x = impl_->slot[who[j]];
    x.conv_at = impl_->host[2 * j];
    x.nan_seen = impl_->host[2 * j + 1];
  }
}

bool KnBatch::running(int slot) const { return impl_->live(impl_->slot[slot]); }

int64_t KnBatch::iterations(int slot) const {
  const Impl::Slot& x = impl_->slot[slot];
  return x.conv_at != INT_MAX ? static_cast<int64_t>(x.conv_at) + 1 : x.t;
}

bool KnBatch::converged(int slot) const { return impl_->slot[slot].conv_at != INT_MAX; }

bool KnBatch::numeric_ok(int slot) const { return impl_->slot[slot].nan_seen == 0; }

void KnBatch::finish(int slot, double* coef, hipStream_t st) {
  Impl::Slot& x = impl_->slot[slot];
  MCNLA_REQUIRE(x.used && !impl_->live(x), MCNLA_ERR_STATE, "KN batch member still running");
  finish_iteration(x.g, x.L, iterations(slot), coef, st);
  x.used = false;
}

void KnBatch::drop(int slot) { impl_->slot[slot].used = false; }

int64_t integrate_gram(const GramIntegrator& g, double* coef, hipStream_t st, bool* converged,
                       bool* numeric_ok) {
  if (!g.wen_yin) {
    int64_t iters = 0;
    bool conv = false, ok = true;
    run_kn(&g, &coef, 1, st, g.graphs, &iters, &conv, &ok);
    if (converged) *converged = conv;
    if (numeric_ok) *numeric_ok = ok;
    return iters;
  }
  Layout L = layout(g.ws, g.N, g.l);
  check_integrator(g, L);
  const int l = g.l;
  const int64_t nl = L.nl;
  const double invN = 1.0 / static_cast<double>(g.N);
  const int ll = l * l;
  const unsigned gp = static_cast<unsigned>(L.nblk_p);
  hipStream_t ls = st;  // where the iteration launches go: a capture stream while recording

  init_iteration(g, L, st);
  const size_t wy_lds = wy_lds_bytes(l);
  for (const void* f : {reinterpret_cast<const void*>(wy_search_kernel),
                        reinterpret_cast<const void*>(wy_tau_kernel)})
    MCNLA_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(wy_lds)));
  // Bg0 = Bs B0 / N and [Dc; Dgc; Hgg] of parity 0.
  bs_product(g, L, L.B[0], L.Bg[0], -1, st);
  hipLaunchKernelGGL(gram_partials_kernel, dim3(gp), dim3(kThreads), 0, st, -1, L.st, nl, l,
                     L.B[0], L.Bg[0], L.P, 3);
  MCNLA_LAUNCHED();
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(ceil_div(3 * ll, kThreads)), dim3(kThreads), 0,
                     st, -1, L.st, L.P, L.nblk_p, 3 * ll, invN, L.D3[0]);
  MCNLA_LAUNCHED();
  hipLaunchKernelGGL(wy_init_scalars_kernel, dim3(1), dim3(64), 0, st, L.st, l, L.D3[0]);
  MCNLA_LAUNCHED();

  int64_t t = 0;
  int64_t chunk = 8;
  int host_conv = INT_MAX;
  int nan_seen = 0;
  while (t < g.maxiter) {
    const int64_t tend = (t + chunk < g.maxiter) ? t + chunk : g.maxiter;
    hipGraphExec_t exec = g.graphs ? g.graphs->find(t, tend) : nullptr;
    if (exec == nullptr && g.graphs) {  // record this chunk once
      if (!g.graphs->cap) MCNLA_HIP(hipStreamCreateWithFlags(&g.graphs->cap, hipStreamNonBlocking));
      ls = g.graphs->cap;
      MCNLA_HIP(hipStreamBeginCapture(ls, hipStreamCaptureModeThreadLocal));
    }
    const int64_t t0 = t;
    const bool recording = exec == nullptr && g.graphs != nullptr;
    if (exec != nullptr) {
      t = tend;
    } else try {
      for (; t < tend; ++t) {
        const int c = static_cast<int>(t & 1), p = c ^ 1;
        const int ti = static_cast<int>(t);
        const int odd = p;  // the reference toggles is_odd before using it: odd = !(t & 1)
        hipLaunchKernelGGL(wy_search_kernel, dim3(1), dim3(kWyThreads), wy_lds, ls, ti, L.st, l,
                           L.D3[c], L.F, L.C);
        MCNLA_LAUNCHED();
        hipLaunchKernelGGL(wy_update_kernel, dim3(wy_update_blocks(nl, l)), dim3(kThreads), 0, ls, ti,
                           L.st, nl, l, invN, L.B[c], L.Bg[c], L.B[p], L.U1, L.BsU1, L.E[c],
                           L.E[p], L.FF[c], L.FF[p], L.F, L.C);
        MCNLA_LAUNCHED();
        bs_product(g, L, L.B[p], L.Bg[p], ti, ls);
        hipLaunchKernelGGL(gram_partials_kernel, dim3(gp), dim3(kThreads), 0, ls, ti, L.st, nl, l,
                           L.B[p], L.Bg[p], L.P, 3);
        MCNLA_LAUNCHED();
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(ceil_div(3 * ll, kThreads)), dim3(kThreads),
                           0, ls, ti, L.st, L.P, L.nblk_p, 3 * ll, invN, L.D3[p]);
        MCNLA_LAUNCHED();
        hipLaunchKernelGGL(wy_check_kernel, dim3(1), dim3(kThreads), 0, ls, ti, L.st, l, g.tol,
                           L.D3[p]);
        MCNLA_LAUNCHED();
        hipLaunchKernelGGL(wy_terms_kernel, dim3(static_cast<unsigned>(L.nblk_t)), dim3(kThreads), 0,
                           ls, ti, L.st, nl, l, invN, odd, L.B[c], L.B[p], L.Bg[c], L.Bg[p], L.U1,
                           L.BsU1, L.D3[p], L.E[c], L.FF[c], L.P);
        MCNLA_LAUNCHED();
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(ceil_div(3 * ll + 2, kThreads)),
                           dim3(kThreads), 0, ls, ti, L.st, L.P, L.nblk_t, 3 * ll + 2, 1.0, L.T);
        MCNLA_LAUNCHED();
        hipLaunchKernelGGL(wy_tau_kernel, dim3(1), dim3(kWyThreads), wy_lds, ls, ti, L.st, l, odd,
                           L.D3[c], L.D3[p], L.F, L.T);
        MCNLA_LAUNCHED();
      }
    } catch (...) {
      if (recording) {  // leave the capture stream usable
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(ls, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        ls = st;
      }
      throw;
    }
    if (recording) {
      hipGraph_t graph = nullptr;
      MCNLA_HIP(hipStreamEndCapture(ls, &graph));
      ls = st;
      hipGraphExec_t made = nullptr;
      const hipError_t e = hipGraphInstantiate(&made, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      MCNLA_HIP(e);
      g.graphs->execs.push_back({{t0, tend}, made});
      exec = made;
    }
    if (exec != nullptr) MCNLA_HIP(hipGraphLaunch(exec, st));
    int fl[2] = {INT_MAX, 0};
    MCNLA_HIP(hipMemcpyAsync(fl, &L.st->converged_at, sizeof(fl), hipMemcpyDeviceToHost, st));
    MCNLA_HIP(hipStreamSynchronize(st));
    host_conv = fl[0];
    nan_seen = fl[1];
    if (host_conv != INT_MAX) break;
    if (chunk < 32) chunk *= 2;
  }
  const int64_t iters = (host_conv != INT_MAX) ? static_cast<int64_t>(host_conv) + 1 : g.maxiter;
  finish_iteration(g, L, iters, coef, st);
  if (converged) *converged = (host_conv != INT_MAX);
  if (numeric_ok) *numeric_ok = (nan_seen == 0);
  return iters;
}

}  // namespace mcnla
