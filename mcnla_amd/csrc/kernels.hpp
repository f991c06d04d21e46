// Host-side launchers of the HIP kernels of libmcnla_amd.
#pragma once

#include <utility>
#include <vector>

#include "common.hpp"

namespace mcnla {

// ------------------------------------------------------------------------------------------------
// FP64 MFMA GEMM (gemm_f64.hip).
enum class GemmShape { kSquare, kTallN32, kTallN64, kWideM32, kSmall };

struct GemmF64Args {
  int64_t M = 0, N = 0, K = 0;
  const void* A = nullptr;  // element (i, k) at A[i + k*lda] if a_mcontig else A[i*lda + k]
  int64_t lda = 0;
  bool a_mcontig = true;
  bool a_f32 = false;
  const void* B = nullptr;  // element (k, j) at B[k*ldb + j] if b_ncontig else B[k + j*ldb]
  int64_t ldb = 0;
  bool b_ncontig = true;
  bool b_f32 = false;
  // Tall FP32-A products only: FP32 MFMA (B rounded to FP32, FP32 sums within a split, FP64
  // across splits) instead of FP64 arithmetic.  Other shapes ignore it.
  bool f32_mfma = false;
  double* C = nullptr;      // row-major, ldc; may be null when only partials are wanted
  int64_t ldc = 0;
  double alpha = 1.0, beta = 0.0;
  int64_t splits = 1;       // split-K factor (requires `partial`)
  int64_t kchunk = 0;       // K range per split (rounded up to the K tile); 0 = K / splits
  double* partial = nullptr;  // splits * M * N scratch for split-K
  bool lower_only = false;  // Gram: compute tiles with tm >= tn only, mirror on reduction
  GemmShape shape = GemmShape::kSquare;
  int64_t batch = 1;        // independent problems; operand b starts at A + b*stride_a, ...
  int64_t stride_a = 0, stride_b = 0, stride_c = 0;
  // Optional device-side gate: when *gate < gate_t the launch does nothing (used by
  // iteration loops that are enqueued ahead of their convergence test).
  const int* gate = nullptr;
  int gate_t = 0;
  // Optional events recorded around the main GEMM launch (not the split-K reduction).
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  // 0: the default kernel for the shape; 1: the alternative one (square shapes: the
  // register-staged 128 x 128 tiles instead of the 256 x 128 DMA tiles; tall shapes: the other
  // tall-skinny kernel -- for an FP32 A on FP32 MFMA register-staged A and both operands
  // streamed by DMA swap (the stream is the default from 1 GiB of A), for an FP64 A the DMA
  // stream instead of A straight into registers, for an FP32 A in FP64 arithmetic A straight
  // into registers instead of register-staged).
  int kernel = 0;
};

void gemm_f64(GemmF64Args p, hipStream_t st);
size_t gemm_f64_partial_elems(int64_t M, int64_t N, int64_t splits, int64_t batch = 1);
// Split-K factor giving about target_blocks workgroups for `tiles` output tiles.
int64_t gemm_f64_splits_for(int64_t tiles, int64_t K, int64_t target_blocks);
// For a DMA-tiled product (M x N, K deep) without split-K: the rows that fill whole rounds of
// the device's workgroups (returned; M when there is nothing to gain) and the split-K factor for
// the remaining rows (*tail_splits, 1 when none), whose last round would otherwise run a few
// full-length workgroups on a mostly idle device.
int64_t gemm_f64_dma_tail(int64_t M, int64_t N, int64_t K, int64_t* tail_splits);
// Compute units of the current device (cached per device).
int compute_units();

// ------------------------------------------------------------------------------------------------
// Mixed-precision sketch (gemm_bf16.hip): Y (FP32, row-major m x nl) = A * Omega with A and
// Omega rounded to BF16 or FP16 (FP16: rows of A scaled by powers of two), FP32 accumulation;
// terms 1: A Omega; 2: A_hi Omega + A_lo Omega; 3: A_hi Omega_hi + A_lo Omega_hi + A_hi Omega_lo.
struct SketchBf16Args {
  int64_t m = 0, n = 0, nl = 0;
  const void* A = nullptr;  // FP32 or FP64 source
  int64_t lda = 0;
  bool a_colmajor = true;
  bool a_f32 = true;
  const double* omega = nullptr;  // n x nl row-major FP64
  float* Y = nullptr;
  double* Yd = nullptr;  // when set, Y is written in FP64 here instead (values are the same)
  int64_t ldy = 0;
  int terms = 1;  // 1, 2 or 3
  int fmt = 0;    // 0: BF16, 1: FP16
  // Optional FP32 scratch for the K-split partials of the rows past the grid's whole rounds
  // (sketch_bf16_tail_bytes); without it those rows run in full-length workgroups.
  float* tail = nullptr;
  size_t tail_bytes = 0;
};
// Packs A and Omega, then runs the GEMM; the optional events bracket the GEMM launch alone.
void sketch_bf16(const SketchBf16Args& a, void* workspace, size_t workspace_bytes, hipStream_t st,
                 hipEvent_t gemm_begin = nullptr, hipEvent_t gemm_end = nullptr);
size_t sketch_bf16_workspace(const SketchBf16Args& a);
// Bytes of SketchBf16Args::tail the split of the last rows takes on this device (0: no split).
size_t sketch_bf16_tail_bytes(const SketchBf16Args& a);

// ------------------------------------------------------------------------------------------------
// Random numbers and synthetic data (rng.hip).
// f16_grid: each value rounded to the nearest FP16 number (kept in FP64).
void omega_philox(double* out, int64_t count, uint64_t seed, hipStream_t st,
                  bool f16_grid = false);
// A = Tm + noise * g (Tm row-major m x n FP64), in A's dtype and storage order.
void synth_finish(int64_t m, int64_t n, const double* Tm, double noise, uint64_t seed, void* A,
                  int64_t lda, bool a_f32, bool a_colmajor, hipStream_t st);
// out[e] = scale * g(e), g the Philox N(0,1) stream of `seed`.
void gaussian_fill(double* out, int64_t count, uint64_t seed, double scale, hipStream_t st);
// Y (m x nl row-major FP64) = the columns idx[0..nl) of A (FP32 or FP64, either order).
void gather_columns(const void* A, int64_t lda, bool a_f32, bool a_colmajor, int64_t m,
                    const int64_t* idx, int64_t nl, double* Y, hipStream_t st);
void synth_lowrank(int64_t m, int64_t n, int64_t r, const double* U, const double* V,
                   const double* sigma, double noise, uint64_t seed, void* A, int64_t lda,
                   bool a_f32, bool a_colmajor, hipStream_t st);
// A = (I - 2uu') [diag(sigma); 0] (I - 2vv') (m x n, n <= m; u, v unit device vectors of m and
// n doubles): singular values exactly sigma, written in A's dtype and order.
void synth_reflected(int64_t m, int64_t n, const double* sigma, const double* u, const double* v,
                     void* A, int64_t lda, bool a_f32, bool a_colmajor, hipStream_t st);

// ------------------------------------------------------------------------------------------------
// Small dense kernels (small.hip).  All FP64, l <= kMaxL.
constexpr int kMaxL = 64;

// Batched symmetric eigen-decomposition (parallel cyclic Jacobi, one workgroup per matrix).
// In: W[b] (n x n row-major, symmetric, stride n*n).  Out: w[b] (n, DESCENDING), V[b] (n x n,
// eigenvectors in columns, each with its largest-magnitude entry positive).
void batched_syev(const double* W, double* w, double* V, int64_t batch, int n, hipStream_t st);

// Orthogonalizer transforms T_i = V_i diag(1/sqrt(w_i)) for every sketch, written into the
// block-diagonal Nl x l stack Tst (row-major; rows i*l..(i+1)*l hold T_i).  Flags non-positive
// eigenvalues in *bad (atomic add).
void orth_transforms(const double* w, const double* V, double* Tst, int64_t N, int l, int* bad,
                     hipStream_t st);

// Copies the N diagonal l x l blocks of an Nl x Nl matrix G into W (N x l x l).
void gram_diag_blocks(const double* G, int64_t ldg, double* W, int64_t N, int l, hipStream_t st);

// Bs = T' G T with T = blockdiag(T_0..T_{N-1}) (Tst as above); Bs and G are Nl x Nl row-major.
void congruence_blockdiag(const double* G, const double* Tst, double* Bs, int64_t N, int l,
                          hipStream_t st);

// T_i = L_i^{-T} with W_i = L_i L_i' (Cholesky) for the N blocks W (N x l x l): Y_i T_i has
// orthonormal columns when W_i = Y_i' Y_i.  L_0 goes to L0 (l x l, lower; may be null).
// Non-positive pivots are flagged in *bad (atomic add).
void cholesky_transforms(const double* W, int64_t N, int l, double* Tst, double* L0, int* bad,
                         hipStream_t st);

// out (m x n, row-major FP64) = A (FP32 or FP64, row- or column-major with leading dimension lda).
void to_rowmajor_f64(const void* A, int64_t lda, bool a_f32, bool colmajor, int64_t m, int64_t n,
                     double* out, hipStream_t st);

// In-place per-sketch apply of the collection: Q_i = Y_i T_i  (Y row-major m x Nl, FP64 or FP32).
void apply_blockdiag(void* Y, bool y_f32, int64_t m, int64_t N, int l, const double* Tst,
                     hipStream_t st);

// ------------------------------------------------------------------------------------------------
// Integrator state machine on the Gram matrix Bs (integrate.hip).
struct IntegratorState;  // device-side scalars (flags, counters, tau, phi, ...)

// HIP graphs of the integrator's chunks of iterations, recorded once and replayed by later solves
// of the same plan.  A chunk [t0, t1) is the same launches with the same arguments every time (the
// iteration index is a kernel argument, the buffers belong to the plan), so its graph can be
// replayed as is.  Owned by the caller; the capture stream is created on first use.
struct IterationGraphs {
  std::vector<std::pair<std::pair<int64_t, int64_t>, hipGraphExec_t>> execs;
  hipStream_t cap = nullptr;
  IterationGraphs() = default;
  IterationGraphs(const IterationGraphs&) = delete;
  IterationGraphs& operator=(const IterationGraphs&) = delete;
  ~IterationGraphs();
  hipGraphExec_t find(int64_t t0, int64_t t1) const {
    for (const auto& e : execs)
      if (e.first.first == t0 && e.first.second == t1) return e.second;
    return nullptr;
  }
};

struct GramIntegrator {
  int64_t N = 0;
  int l = 0;
  bool wen_yin = false;
  int64_t maxiter = 256;
  double tol = 1e-4;
  // device buffers (owned by the caller, sized by integrator_workspace_bytes)
  double* Bs = nullptr;   // Nl x Nl
  char* ws = nullptr;
  size_t ws_bytes = 0;
  IterationGraphs* graphs = nullptr;  // record/replay the iteration chunks as HIP graphs
};
size_t integrator_workspace_bytes(int64_t N, int l);
// Runs the whole iteration on the device.  On return (stream-ordered) the Nl x l coefficient
// matrix M with Qbar = Qs * M is in `coef` (Qs = the orthonormal collection whose Gram is Bs).
// Returns the iteration count (this call synchronizes the stream once per chunk of iterations).
int64_t integrate_gram(const GramIntegrator& g, double* coef, hipStream_t st, bool* converged,
                       bool* numeric_ok);

// KN integrations of up to kMaxKnGroup solves (wen_yin false; any N, l, tol, maxiter each) run in
// lockstep on one stream: each iteration is one launch per kernel for the whole group (the
// Bs product: one per tile-width class, l <= 32 and l > 32), and the host reads the group's
// convergence flags once per chunk of iterations.  Member i's coefficients, iteration count and
// flags are bit-identical to integrate_gram(gs[i]) alone.  `graphs` of the members are not used.
constexpr int kMaxKnGroup = 16;
void integrate_gram_group(const GramIntegrator* gs, double* const* coefs, int count,
                          hipStream_t st, int64_t* iterations, bool* converged, bool* numeric_ok);

// KN integrations that join and leave a running group between chunks of iterations (the serving
// loop's iteration-level batching).  Each member keeps its own iteration index and state, so its
// coefficients, iteration count and flags are bit-identical to integrate_gram alone, whatever it
// shares launches with and whenever it joined.  Not thread-safe; one stream at a time.
class KnBatch {
 public:
  // `flags`: device memory for 2 kMaxKnGroup ints (the per-chunk read-back of the flags).
  explicit KnBatch(int* flags);
  ~KnBatch();
  KnBatch(const KnBatch&) = delete;
  KnBatch& operator=(const KnBatch&) = delete;
  int free_slots() const;
  // Takes a KN integration (its Bs complete in stream order on `st`) and enqueues its initial
  // state; returns its slot.
  int add(const GramIntegrator& g, hipStream_t st);
  // Enqueues up to `iters` iterations of each running member, then the read-back of every
  // member's flags, and synchronizes `st`.
  void advance(int iters, hipStream_t st);
  bool running(int slot) const;  // neither converged nor at its maxiter
  int64_t iterations(int slot) const;
  bool converged(int slot) const;
  bool numeric_ok(int slot) const;
  // Enqueues the member's final coefficients into `coef` (Nl x l) and releases its slot.
  void finish(int slot, double* coef, hipStream_t st);
  // Releases a slot without finishing.
  void drop(int slot);

 private:
  struct Impl;
  Impl* impl_;
};

// ------------------------------------------------------------------------------------------------
// Non-iterative integrators on Bs (combine.hip): the pairwise reduction tree and the extrinsic
// mean.  Both leave the Nl x l coefficients with Qbar = Qs * coef in `coef`.
size_t combine_workspace_bytes(bool extrinsic, int64_t N, int l);
void integrate_reduction(const double* Bs, int64_t N, int l, double* coef, char* ws,
                         size_t ws_bytes, hipStream_t st);
void integrate_extrinsic(const double* Bs, int64_t N, int l, double* coef, char* ws,
                         size_t ws_bytes, hipStream_t st);

// ------------------------------------------------------------------------------------------------
// Former helpers (former.hip).
// Given Z = A' Qbar (n x l row-major) and Wg = Z'Z (l x l): eig, S (k), the l x k rotation
// Wk (row-major), and, if V != null, V = Z Wk diag(1/S) (n x k row-major).
// symmetric: Wg = Qbar' A Qbar of a symmetric A; S and Wk take the k largest eigenpairs in
// ascending order and S holds the eigenvalues themselves.
void former_small(const double* Wg, int l, int k, double* eigw, double* eigv, double* Wk,
                  double* S, hipStream_t st, bool symmetric = false);
void scale_columns(double* X, int64_t rows, int k, const double* S, bool inverse, hipStream_t st);
void column_norms(const double* X, int64_t rows, int k, double* out, hipStream_t st);

}  // namespace mcnla
