// Batched small-matrix kernels of the orthogonalizer and the formers (FP64).
#include "common.hpp"
#include "kernels.hpp"
#include "small.hpp"

namespace mcnla {

namespace {

constexpr int kSmallThreads = 256;
constexpr int kLds = kMaxL * (kMaxL + 1);
// A Jacobi round applies up to (l/2)^2 + (l/2) l = 3072 work items at l = 64: 1024 threads take
// three each (256 threads took twelve, one after another, each waiting on its LDS reads).
constexpr int kSyevThreads = 1024;

// One workgroup per matrix: W[b] -> (w[b] descending, V[b]).
__global__ __launch_bounds__(kSyevThreads) void batched_syev_kernel(const double* __restrict__ W,
                                                                   double* __restrict__ w,
                                                                   double* __restrict__ V,
                                                                   int n) {
  __shared__ double a[kLds];
  __shared__ double v[kLds];
  __shared__ double cs[2 * kMaxL];
  __shared__ int pq[2 * kMaxL];
  __shared__ double red[2 * (kSyevThreads / kWave)];
  __shared__ int ord[kMaxL];
  const int64_t b = blockIdx.x;
  const double* Wb = W + b * n * n;
  const int ld = lds_ld(n);
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    // symmetrize from the lower triangle (i >= j)
    a[i * ld + j] = (i >= j) ? Wb[i * n + j] : Wb[j * n + i];
  }
  __syncthreads();
  lds_jacobi(n, a, v, true, cs, pq, red);
  lds_eig_sorted(n, a, v, w + b * n, V + b * n * n, ord);
}

// T_i = V_i diag(1/sqrt(w_i)) into rows [i*l, (i+1)*l) of the Nl x l stack.
__global__ void orth_transforms_kernel(const double* __restrict__ w, const double* __restrict__ V,
                                       double* __restrict__ T, int64_t N, int l, int* bad) {
  const int64_t total = N * l * l;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / (l * l);
    const int r = (e / l) % l, c = e % l;
    const double ev = w[b * l + c];
    if (r == 0) {
      // an eigenvalue at or below l eps of the block's largest is zero at working precision:
      // the sketch does not have full column rank (or holds NaN / Inf)
      double wmax = 0.0;
      for (int q = 0; q < l; ++q) wmax = fmax(wmax, w[b * l + q]);
      if (!(ev > l * 0x1p-52 * wmax)) atomicAdd(bad, 1);
    }
    T[e] = V[e] / sqrt(ev);
  }
}

__global__ void gram_diag_blocks_kernel(const double* __restrict__ G, int64_t ldg,
                                        double* __restrict__ W, int64_t N, int l) {
  const int64_t total = N * l * l;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / (l * l);
    const int r = (e / l) % l, c = e % l;
    W[e] = G[(b * l + r) * ldg + b * l + c];
  }
}

// Bs = blockdiag(T)' G blockdiag(T), block (I, J) = T_I' G_IJ T_J for I >= J: one workgroup per
// block pair, G_IJ, T_I, T_J and H = G_IJ T_J in LDS (dynamic, 4 np (np + 1) doubles, np =
// pad16(l), zero-padded) and both products on f64 MFMA tiles.  Bs is symmetric in exact
// arithmetic; the entries with global row >= column are computed and written to both halves, so
// that it is exactly symmetric (its rows can then stand in for its columns).
__global__ __launch_bounds__(256) void congruence_kernel(const double* __restrict__ G,
                                                        const double* __restrict__ T,
                                                        double* __restrict__ Bs, int64_t N,
                                                        int l) {
  extern __shared__ double smem[];
  const int np = pad16(l), ld = np + 1, nb = np * ld;
  double* sG = smem;
  double* sTi = sG + nb;
  double* sTj = sTi + nb;
  double* sH = sTj + nb;
  const int64_t nl = N * l;
  // block pair p -> (I, J), I >= J, p = I (I + 1) / 2 + J
  const int64_t p = blockIdx.x;
  int64_t I = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(p) + 1.0) - 1.0) / 2.0);
  while (I * (I + 1) / 2 > p) --I;
  while ((I + 1) * (I + 2) / 2 <= p) ++I;
  const int64_t J = p - I * (I + 1) / 2;
  const int ll = l * l;
  for (int e = threadIdx.x; e < np * np; e += blockDim.x) {
    const int a = e / np, b = e % np;
    const bool in = a < l && b < l;
    sG[a * ld + b] = in ? G[(I * l + a) * nl + J * l + b] : 0.0;
    sTi[a * ld + b] = in ? T[I * ll + a * l + b] : 0.0;
    sTj[a * ld + b] = in ? T[J * ll + a * l + b] : 0.0;
  }
  __syncthreads();
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int nt1 = np / 16, tiles = nt1 * nt1, nw = blockDim.x / kWave;
  for (int t = wave; t < tiles; t += nw) {  // H = G_IJ T_J
    const int i0 = (t / nt1) * 16, j0 = (t % nt1) * 16;
    const d4 acc = lds_mfma_tile<false, false>(np, ld, sG, sTj, i0, j0);
#pragma unroll
    for (int r = 0; r < 4; ++r) sH[(i0 + lane / 16 + 4 * r) * ld + j0 + lane % 16] = acc[r];
  }
  __syncthreads();
  for (int t = wave; t < tiles; t += nw) {  // T_I' H
    const int i0 = (t / nt1) * 16, j0 = (t % nt1) * 16;
    const d4 acc = lds_mfma_tile<true, false>(np, ld, sTi, sH, i0, j0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = i0 + lane / 16 + 4 * q, c = j0 + lane % 16;
      if (r >= l || c >= l || (I == J && r < c)) continue;
      Bs[(I * l + r) * nl + J * l + c] = acc[q];
      Bs[(J * l + c) * nl + I * l + r] = acc[q];
    }
  }
}

// T_i = L_i^{-T} from the Cholesky factorization W_i = L_i L_i' (one workgroup per block), so
// that Y_i T_i has orthonormal columns when W_i = Y_i' Y_i.  L_0 is also written to L0 (lower
// triangle, zeros above).  A pivot that is not above l eps of W_i's largest diagonal entry flags
// *bad.
__global__ __launch_bounds__(256) void cholesky_transforms_kernel(const double* __restrict__ W,
                                                                 double* __restrict__ T,
                                                                 double* __restrict__ L0, int l,
                                                                 int* bad) {
  __shared__ double a[kLds];
  __shared__ double x[kLds];
  const int ld = lds_ld(l);
  const int64_t b = blockIdx.x;
  const double* Wb = W + b * l * l;
  const int tid = threadIdx.x;
  for (int e = tid; e < l * l; e += blockDim.x) {
    const int i = e / l, j = e % l;
    a[i * ld + j] = (i >= j) ? Wb[i * l + j] : 0.0;
    x[i * ld + j] = 0.0;
  }
  __syncthreads();
  // A pivot at or below l eps of the largest diagonal entry is zero at working precision: the
  // sketch does not have full column rank (or W holds NaN / Inf, which fails the test too).
  double wmax = 0.0;
  for (int i = 0; i < l; ++i) wmax = fmax(wmax, Wb[i * l + i]);
  const double floor_piv = l * 0x1p-52 * wmax;
  bool fail = false;
  for (int j = 0; j < l; ++j) {
    const double d = a[j * ld + j];
    const double piv = (d > 0.0) ? sqrt(d) : 1.0;
    fail = fail || !(d > floor_piv);
    __syncthreads();  // every thread has read the pivot before it is overwritten
    if (tid == 0) a[j * ld + j] = piv;
    for (int r = j + 1 + tid; r < l; r += blockDim.x) a[r * ld + j] /= piv;
    __syncthreads();
    const int w = l - 1 - j;
    for (int e = tid; e < w * w; e += blockDim.x) {
      const int r = j + 1 + e / w, c = j + 1 + e % w;
      if (c <= r) a[r * ld + c] -= a[r * ld + j] * a[c * ld + j];
    }
    __syncthreads();
  }
  // X = L^{-1}: thread c solves L x = e_c by forward substitution (column c of X in LDS)
  for (int c = tid; c < l; c += blockDim.x) {
    x[c * ld + c] = 1.0 / a[c * ld + c];
    for (int r = c + 1; r < l; ++r) {
      double s = 0.0;
      for (int q = c; q < r; ++q) s += a[r * ld + q] * x[q * ld + c];
      x[r * ld + c] = -s / a[r * ld + r];
    }
  }
  __syncthreads();
  double* Tb = T + b * l * l;
  for (int e = tid; e < l * l; e += blockDim.x) {
    const int i = e / l, j = e % l;
    Tb[e] = x[j * ld + i];  // T = X'
    if (b == 0 && L0 != nullptr) L0[e] = (i >= j) ? a[i * ld + j] : 0.0;
  }
  if (fail && tid == 0) atomicAdd(bad, 1);
}

// out (m x n, row-major FP64, ld n) = A (FP32 or FP64, row- or column-major, lda), through
// 32 x 32 LDS tiles so that reads and writes are both coalesced.
template <typename TS, bool COLMAJOR>
__global__ __launch_bounds__(256) void to_rowmajor_f64_kernel(const TS* __restrict__ A, int64_t lda,
                                                             int64_t m, int64_t n,
                                                             double* __restrict__ out) {
  __shared__ double tile[32][33];
  const int64_t r0 = static_cast<int64_t>(blockIdx.y) * 32, c0 = static_cast<int64_t>(blockIdx.x) * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (COLMAJOR) {  // column c0 + ty + 8q, rows r0 + tx contiguous
      const int64_t r = r0 + tx, c = c0 + ty + 8 * q;
      tile[tx][ty + 8 * q] = (r < m && c < n) ? static_cast<double>(A[r + c * lda]) : 0.0;
    } else {
      const int64_t r = r0 + ty + 8 * q, c = c0 + tx;
      tile[ty + 8 * q][tx] = (r < m && c < n) ? static_cast<double>(A[r * lda + c]) : 0.0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t r = r0 + ty + 8 * q, c = c0 + tx;
    if (r < m && c < n) out[r * n + c] = tile[ty + 8 * q][tx];
  }
}

// Q_i = Y_i T_i in place; one thread per (row, sketch).
template <typename T>
__global__ void apply_blockdiag_kernel(T* __restrict__ Y, int64_t m, int64_t N, int l,
                                       const double* __restrict__ Tst) {
  const int64_t total = m * N;
  const int64_t nl = N * l;
  double y[kMaxL];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / N, b = e % N;
    T* row = Y + r * nl + b * l;
    for (int q = 0; q < l; ++q) y[q] = static_cast<double>(row[q]);
    const double* t = Tst + b * l * l;
    for (int c = 0; c < l; ++c) {
      double acc = 0.0;
      for (int q = 0; q < l; ++q) acc += y[q] * t[q * l + c];
      row[c] = static_cast<T>(acc);
    }
  }
}

// Former: eig of Wg (l x l) -> eigw (descending), eigv; Wk = eigv[:, :k]; S = sqrt(eigw[:k]).
// symmetric (Wg = Qbar' A Qbar of a symmetric A): the k largest eigenpairs in ascending order,
// S = the eigenvalues themselves (row_block_symmetric_former.hpp:65-66,127-130).
// One l x l problem on one workgroup: 16 waves share each Jacobi round's work.
constexpr int kFormerThreads = 1024;

__global__ __launch_bounds__(kFormerThreads) void former_small_kernel(const double* __restrict__ Wg,
                                                                    int l, int k, bool symmetric,
                                                                    double* __restrict__ eigw,
                                                                    double* __restrict__ eigv,
                                                                    double* __restrict__ Wk,
                                                                    double* __restrict__ S) {
  __shared__ double a[kLds];
  __shared__ double v[kLds];
  __shared__ double cs[2 * kMaxL];
  __shared__ int pq[2 * kMaxL];
  __shared__ double red[2 * (kFormerThreads / kWave)];
  __shared__ int ord[kMaxL];
  const int ld = lds_ld(l);
  for (int e = threadIdx.x; e < l * l; e += blockDim.x) {
    const int i = e / l, j = e % l;
    a[i * ld + j] = 0.5 * (Wg[i * l + j] + Wg[j * l + i]);
  }
  __syncthreads();
  lds_jacobi(l, a, v, true, cs, pq, red);
  lds_eig_sorted(l, a, v, eigw, eigv, ord);
  for (int e = threadIdx.x; e < l * k; e += blockDim.x) {
    const int i = e / k, j = e % k;
    Wk[e] = eigv[i * l + (symmetric ? k - 1 - j : j)];
  }
  for (int j = threadIdx.x; j < k; j += blockDim.x)
    S[j] = symmetric ? eigw[k - 1 - j] : sqrt(fmax(eigw[j], 0.0));
}

__global__ void scale_columns_kernel(double* __restrict__ X, int64_t rows, int k,
                                     const double* __restrict__ S, bool inverse) {
  const int64_t total = rows * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double s = S[e % k];
    X[e] = inverse ? X[e] / s : X[e] * s;
  }
}

// out[j] = ||X[:, j]||_2 for a row-major rows x k matrix; one workgroup per column.
__global__ void column_norms_kernel(const double* __restrict__ X, int64_t rows, int k,
                                    double* __restrict__ out) {
  __shared__ double red[kSmallThreads / kWave];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) {
    const double x = X[r * k + j];
    acc += x * x;
  }
  acc = wave_sum(acc);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s += red[w];
    out[j] = sqrt(s);
  }
}

unsigned grid_for(int64_t total, int threads = 256, int64_t cap = 16384) {
  int64_t b = ceil_div(total, threads);
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

}  // namespace

void batched_syev(const double* W, double* w, double* V, int64_t batch, int n, hipStream_t st) {
  MCNLA_REQUIRE(n >= 1 && n <= kMaxL, MCNLA_ERR_INVALID, "eigen problem larger than 64");
  if (batch <= 0) return;
  hipLaunchKernelGGL(batched_syev_kernel, dim3(static_cast<unsigned>(batch)), dim3(kSyevThreads),
                     0, st, W, w, V, n);
  MCNLA_LAUNCHED();
}

void orth_transforms(const double* w, const double* V, double* Tst, int64_t N, int l, int* bad,
                     hipStream_t st) {
  hipLaunchKernelGGL(orth_transforms_kernel, dim3(grid_for(N * l * l)), dim3(256), 0, st, w, V, Tst,
                     N, l, bad);
  MCNLA_LAUNCHED();
}

void gram_diag_blocks(const double* G, int64_t ldg, double* W, int64_t N, int l, hipStream_t st) {
  hipLaunchKernelGGL(gram_diag_blocks_kernel, dim3(grid_for(N * l * l)), dim3(256), 0, st, G, ldg,
                     W, N, l);
  MCNLA_LAUNCHED();
}

void congruence_blockdiag(const double* G, const double* Tst, double* Bs, int64_t N, int l,
                          hipStream_t st) {
  MCNLA_REQUIRE(l >= 1 && l <= kMaxL, MCNLA_ERR_INVALID, "l > 64");
  const int64_t pairs = N * (N + 1) / 2;
  if (pairs == 0) return;
  const size_t lds = static_cast<size_t>(4) * pad16(l) * (pad16(l) + 1) * sizeof(double);
  hipLaunchKernelGGL(congruence_kernel, dim3(static_cast<unsigned>(pairs)), dim3(256), lds, st, G,
                     Tst, Bs, N, l);
  MCNLA_LAUNCHED();
}

void to_rowmajor_f64(const void* A, int64_t lda, bool a_f32, bool colmajor, int64_t m, int64_t n,
                     double* out, hipStream_t st) {
  if (m == 0 || n == 0) return;
  MCNLA_REQUIRE(ceil_div(m, 32) < 65536, MCNLA_ERR_INVALID, "too many rows for the converter");
  const dim3 grid(static_cast<unsigned>(ceil_div(n, 32)), static_cast<unsigned>(ceil_div(m, 32)));
  if (a_f32) {
    if (colmajor)
      hipLaunchKernelGGL((to_rowmajor_f64_kernel<float, true>), grid, dim3(256), 0, st,
                         static_cast<const float*>(A), lda, m, n, out);
    else
      hipLaunchKernelGGL((to_rowmajor_f64_kernel<float, false>), grid, dim3(256), 0, st,
                         static_cast<const float*>(A), lda, m, n, out);
  } else {
    if (colmajor)
      hipLaunchKernelGGL((to_rowmajor_f64_kernel<double, true>), grid, dim3(256), 0, st,
                         static_cast<const double*>(A), lda, m, n, out);
    else
      hipLaunchKernelGGL((to_rowmajor_f64_kernel<double, false>), grid, dim3(256), 0, st,
                         static_cast<const double*>(A), lda, m, n, out);
  }
  MCNLA_LAUNCHED();
}

void cholesky_transforms(const double* W, int64_t N, int l, double* Tst, double* L0, int* bad,
                         hipStream_t st) {
  MCNLA_REQUIRE(l >= 1 && l <= kMaxL, MCNLA_ERR_INVALID, "l > 64");
  if (N <= 0) return;
  hipLaunchKernelGGL(cholesky_transforms_kernel, dim3(static_cast<unsigned>(N)), dim3(256), 0, st,
                     W, Tst, L0, l, bad);
  MCNLA_LAUNCHED();
}

void apply_blockdiag(void* Y, bool y_f32, int64_t m, int64_t N, int l, const double* Tst,
                     hipStream_t st) {
  MCNLA_REQUIRE(l <= kMaxL, MCNLA_ERR_INVALID, "l > 64");
  if (y_f32) {
    hipLaunchKernelGGL(apply_blockdiag_kernel<float>, dim3(grid_for(m * N)), dim3(256), 0, st,
                       static_cast<float*>(Y), m, N, l, Tst);
  } else {
    hipLaunchKernelGGL(apply_blockdiag_kernel<double>, dim3(grid_for(m * N)), dim3(256), 0, st,
                       static_cast<double*>(Y), m, N, l, Tst);
  }
  MCNLA_LAUNCHED();
}

void former_small(const double* Wg, int l, int k, double* eigw, double* eigv, double* Wk,
                  double* S, hipStream_t st, bool symmetric) {
  MCNLA_REQUIRE(l >= 1 && l <= kMaxL && k >= 1 && k <= l, MCNLA_ERR_INVALID, "former sizes");
  hipLaunchKernelGGL(former_small_kernel, dim3(1), dim3(kFormerThreads), 0, st, Wg, l, k, symmetric,
                     eigw, eigv, Wk, S);
  MCNLA_LAUNCHED();
}

void scale_columns(double* X, int64_t rows, int k, const double* S, bool inverse,
                   hipStream_t st) {
  hipLaunchKernelGGL(scale_columns_kernel, dim3(grid_for(rows * k)), dim3(256), 0, st, X, rows, k,
                     S, inverse);
  MCNLA_LAUNCHED();
}

void column_norms(const double* X, int64_t rows, int k, double* out, hipStream_t st) {
  hipLaunchKernelGGL(column_norms_kernel, dim3(static_cast<unsigned>(k)), dim3(kSmallThreads), 0,
                     st, X, rows, k, out);
  MCNLA_LAUNCHED();
}

}  // namespace mcnla
