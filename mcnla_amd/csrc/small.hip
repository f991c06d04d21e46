// Batched small-matrix kernels of the orthogonalizer and the formers (FP64).
#include "common.hpp"
#include "kernels.hpp"
#include "small.hpp"

namespace mcnla {

namespace {

constexpr int kSmallThreads = 256;
constexpr int kLds = kMaxL * (kMaxL + 1);

// One workgroup per matrix: W[b] -> (w[b] descending, V[b]).
__global__ __launch_bounds__(kSmallThreads) void batched_syev_kernel(const double* __restrict__ W,
                                                                    double* __restrict__ w,
                                                                    double* __restrict__ V,
                                                                    int n) {
  __shared__ double a[kLds];
  __shared__ double v[kLds];
  __shared__ double cs[2 * kMaxL];
  __shared__ int pq[2 * kMaxL];
  __shared__ double red[2 * (kSmallThreads / kWave)];
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
    if (r == 0 && !(ev > 0.0)) atomicAdd(bad, 1);
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

// H = G * blockdiag(T): H[r][c] = sum_e G[r][blk(c)*l + e] * T_blk(c)[e][c % l]
__global__ void right_blockdiag_kernel(const double* __restrict__ G, const double* __restrict__ T,
                                       double* __restrict__ H, int64_t nl, int l) {
  const int64_t total = nl * nl;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / nl, c = e % nl;
    const int64_t blk = c / l;
    const int cc = c % l;
    const double* g = G + r * nl + blk * l;
    const double* t = T + blk * l * l + cc;
    double acc = 0.0;
    for (int q = 0; q < l; ++q) acc += g[q] * t[q * l];
    H[e] = acc;
  }
}

// Bs = blockdiag(T)' * H: Bs[r][c] = sum_e T_blk(r)[e][r % l] * H[blk(r)*l + e][c].  Bs is
// symmetric in exact arithmetic; the entry with r >= c is computed and written to both halves so
// that it is exactly symmetric (its rows can then stand in for its columns).
__global__ void left_blockdiag_t_kernel(const double* __restrict__ H, const double* __restrict__ T,
                                        double* __restrict__ Bs, int64_t nl, int l) {
  const int64_t total = nl * nl;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / nl, c = e % nl;
    if (r < c) continue;
    const int64_t blk = r / l;
    const int rr = r % l;
    const double* h = H + blk * l * nl + c;
    const double* t = T + blk * l * l + rr;
    double acc = 0.0;
    for (int q = 0; q < l; ++q) acc += t[q * l] * h[q * nl];
    Bs[e] = acc;
    Bs[c * nl + r] = acc;
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
__global__ __launch_bounds__(kSmallThreads) void former_small_kernel(const double* __restrict__ Wg,
                                                                    int l, int k, bool symmetric,
                                                                    double* __restrict__ eigw,
                                                                    double* __restrict__ eigv,
                                                                    double* __restrict__ Wk,
                                                                    double* __restrict__ S) {
  __shared__ double a[kLds];
  __shared__ double v[kLds];
  __shared__ double cs[2 * kMaxL];
  __shared__ int pq[2 * kMaxL];
  __shared__ double red[2 * (kSmallThreads / kWave)];
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
  hipLaunchKernelGGL(batched_syev_kernel, dim3(static_cast<unsigned>(batch)), dim3(kSmallThreads),
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

void congruence_blockdiag(const double* G, const double* Tst, double* Bs, double* tmp, int64_t N,
                          int l, hipStream_t st) {
  const int64_t nl = N * l;
  hipLaunchKernelGGL(right_blockdiag_kernel, dim3(grid_for(nl * nl)), dim3(256), 0, st, G, Tst, tmp,
                     nl, l);
  MCNLA_LAUNCHED();
  hipLaunchKernelGGL(left_blockdiag_t_kernel, dim3(grid_for(nl * nl)), dim3(256), 0, st, tmp, Tst,
                     Bs, nl, l);
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
  hipLaunchKernelGGL(former_small_kernel, dim3(1), dim3(kSmallThreads), 0, st, Wg, l, k, symmetric,
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
