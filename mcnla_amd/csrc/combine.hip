// The non-iterative integrators on the Gram matrix Bs = Qs'Qs (Nl x Nl, FP64): the pairwise
// reduction tree (row_block_reduction_integrator.hpp:86-120) and the extrinsic mean
// (row_block_extrinsic_mean_integrator.hpp:105-162).  As for KN and WY, the sketches are only
// touched through Bs and the result is a coefficient matrix coef (Nl x l) with Qbar = Qs * coef.
#include "common.hpp"
#include "kernels.hpp"
#include "small.hpp"

namespace mcnla {

namespace {

constexpr int kThreads = 256;

unsigned grid_for(int64_t total) {
  int64_t b = ceil_div(total, kThreads);
  if (b > 8192) b = 8192;
  return static_cast<unsigned>(b < 1 ? 1 : b);
}

// C[i] (Nl x l) = the selector of sketch i: rows i*l .. i*l + l - 1 hold the identity.
__global__ void selectors_kernel(double* __restrict__ C, int64_t N, int64_t nl, int l) {
  const int64_t total = N * nl * l;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / (nl * l), r = (e / l) % nl;
    const int c = static_cast<int>(e % l);
    C[e] = (r == i * l + c) ? 1.0 : 0.0;
  }
}

// Sorts the eigenpairs left by lds_jacobi (diagonal of a, columns of v; LDS, ld = n + 1) into
// descending order inside LDS: w[j], vs (columns), each vector's largest-magnitude entry positive.
__device__ void lds_sort_desc(int n, const double* a, const double* v, double* w, double* vs,
                              int* ord) {
  const int ld = lds_ld(n);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double di = isnan(a[i * ld + i]) ? -INFINITY : a[i * ld + i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = isnan(a[j * ld + j]) ? -INFINITY : a[j * ld + j];
      if (dj > di || (dj == di && j < i)) ++r;
    }
    ord[r] = i;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const int src = ord[j];
    w[j] = a[src * ld + src];
    int big = 0;
    for (int i = 1; i < n; ++i)
      if (fabs(v[i * ld + src]) > fabs(v[big * ld + src])) big = i;
    const double sg = v[big * ld + src] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) vs[i * ld + j] = sg * v[i * ld + src];
  }
  __syncthreads();
}

// One tree level, one workgroup per pair i < h: with B = C_i' Bs C_{i+h} (given, l x l),
// svd(B) = W S T' (T and S^2 from eig(B'B), W = B T S^-1), then in place
// C_i = (C_i W + C_{i+h} T) diag(1 / sqrt(2 (1 + s))).
__global__ __launch_bounds__(kThreads) void reduction_pair_kernel(const double* __restrict__ Bt,
                                                                 double* __restrict__ C,
                                                                 int64_t nl, int l, int64_t h) {
  constexpr int kLd = kMaxL + 1;
  constexpr int kRowsChunk = 64;
  __shared__ double sb[kMaxL * kLd];   // B, later rows of C_i
  __shared__ double sa[kMaxL * kLd];   // B'B -> eigenvalues; later rows of C_{i+h}
  __shared__ double sv[kMaxL * kLd];   // Jacobi vectors, then W
  __shared__ double st[kMaxL * kLd];   // T (sorted)
  __shared__ double w[kMaxL], d[kMaxL], cs[2 * kMaxL], red[8];
  __shared__ int pq[2 * kMaxL], ord[kMaxL];
  static_assert(kRowsChunk * kMaxL <= kMaxL * kLd, "row chunk fits a buffer");
  const int64_t i = blockIdx.x;
  const int ld = lds_ld(l), tid = threadIdx.x;
  const double* Bi = Bt + i * l * l;
  for (int e = tid; e < l * l; e += kThreads) sb[(e / l) * ld + e % l] = Bi[e];
  __syncthreads();
  for (int e = tid; e < l * l; e += kThreads) {  // B'B
    const int r = e / l, c = e % l;
    double acc = 0.0;
    for (int q = 0; q < l; ++q) acc += sb[q * ld + r] * sb[q * ld + c];
    sa[r * ld + c] = acc;
  }
  __syncthreads();
  for (int e = tid; e < l * l; e += kThreads) {  // exact symmetry for Jacobi
    const int r = e / l, c = e % l;
    if (r > c) sa[r * ld + c] = sa[c * ld + r];
  }
  __syncthreads();
  lds_jacobi(l, sa, sv, true, cs, pq, red);
  lds_sort_desc(l, sa, sv, w, st, ord);
  for (int j = tid; j < l; j += kThreads) {
    const double s = sqrt(fmax(w[j], 0.0));
    w[j] = s;
    d[j] = 1.0 / sqrt(2.0 * (1.0 + s));
  }
  __syncthreads();
  for (int e = tid; e < l * l; e += kThreads) {  // W = B T / s
    const int r = e / l, c = e % l;
    double acc = 0.0;
    for (int q = 0; q < l; ++q) acc += sb[r * ld + q] * st[q * ld + c];
    sv[r * ld + c] = w[c] > 0.0 ? acc / w[c] : 0.0;
  }
  __syncthreads();
  double* Ci = C + i * nl * l;
  const double* Ch = C + (i + h) * nl * l;
  for (int64_t r0 = 0; r0 < nl; r0 += kRowsChunk) {
    const int rows = static_cast<int>(nl - r0 < kRowsChunk ? nl - r0 : kRowsChunk);
    for (int e = tid; e < rows * l; e += kThreads) {
      sb[e] = Ci[r0 * l + e];
      sa[e] = Ch[r0 * l + e];
    }
    __syncthreads();
    for (int e = tid; e < rows * l; e += kThreads) {
      const int r = e / l, c = e % l;
      double acc = 0.0;
      for (int q = 0; q < l; ++q) acc += sb[r * l + q] * sv[q * ld + c] + sa[r * l + q] * st[q * ld + c];
      Ci[r0 * l + e] = acc * d[c];
    }
    __syncthreads();
  }
}

// Extrinsic mean, one workgroup per sketch i: with V_i the eigenvectors (columns, descending) of
// G_i = B_i B_i', flip column j of V_i when V_0(:,j)' (Q_0'Q_i) V_i(:,j) < 0 (Q_0'Q_i is the
// (0, i) block of Bs), and write it into rows i*l .. i*l + l - 1 of M (Nl x l).
__global__ __launch_bounds__(kThreads) void extrinsic_align_kernel(const double* __restrict__ Bs,
                                                                  const double* __restrict__ V,
                                                                  int64_t nl, int l,
                                                                  double* __restrict__ M) {
  constexpr int kLd = kMaxL + 1;
  __shared__ double s0[kMaxL * kLd], si[kMaxL * kLd], sb[kMaxL * kLd], sx[kMaxL * kLd];
  __shared__ double sgn[kMaxL];
  const int64_t i = blockIdx.x;
  const int ld = lds_ld(l), tid = threadIdx.x;
  for (int e = tid; e < l * l; e += kThreads) {
    const int r = e / l, c = e % l;
    s0[r * ld + c] = V[e];
    si[r * ld + c] = V[i * l * l + e];
    sb[r * ld + c] = Bs[r * nl + i * l + c];
  }
  __syncthreads();
  for (int e = tid; e < l * l; e += kThreads) {  // X = (Q_0'Q_i) V_i
    const int r = e / l, c = e % l;
    double acc = 0.0;
    for (int q = 0; q < l; ++q) acc += sb[r * ld + q] * si[q * ld + c];
    sx[r * ld + c] = acc;
  }
  __syncthreads();
  for (int j = tid; j < l; j += kThreads) {
    double acc = 0.0;
    for (int a = 0; a < l; ++a) acc += s0[a * ld + j] * sx[a * ld + j];
    sgn[j] = acc < 0.0 ? -1.0 : 1.0;
  }
  __syncthreads();
  for (int e = tid; e < l * l; e += kThreads) {
    const int r = e / l, c = e % l;
    M[(i * l + r) * l + c] = sgn[c] * si[r * ld + c];
  }
}

// out = X W diag(scale) for X rows x l, W l x l (row-major), scale[j] = 1 / sqrt(w[j]).
__global__ __launch_bounds__(kThreads) void rmul_invsqrt_kernel(const double* __restrict__ X,
                                                               int64_t rows, int l,
                                                               const double* __restrict__ W,
                                                               const double* __restrict__ w,
                                                               double* __restrict__ out) {
  __shared__ double sW[kMaxL * kMaxL];
  __shared__ double sc[kMaxL];
  for (int e = threadIdx.x; e < l * l; e += kThreads) sW[e] = W[e];
  for (int j = threadIdx.x; j < l; j += kThreads) sc[j] = w[j] > 0.0 ? 1.0 / sqrt(w[j]) : 0.0;
  __syncthreads();
  const int64_t total = rows * l;
  for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kThreads) {
    const int64_t r = e / l;
    const int c = static_cast<int>(e % l);
    double acc = 0.0;
    for (int q = 0; q < l; ++q) acc += X[r * l + q] * sW[q * l + c];
    out[e] = acc * sc[c];
  }
}

struct Carve {
  char* base;
  size_t off = 0;
  double* take(size_t elems) {
    double* p = reinterpret_cast<double*>(base ? base + off : nullptr);
    off += ((elems * sizeof(double) + 255) / 256) * 256;
    return p;
  }
};

int64_t small_splits(int64_t batch, int64_t K) { return gemm_f64_splits_for(batch, K, 1024); }

// Batched l x l products P_b = op(A_b) B_b with K = Nl rows, split over K.
void small_products(GemmF64Args g, int64_t batch, double* partial, hipStream_t st) {
  g.batch = batch;
  g.shape = g.M <= 32 ? GemmShape::kSmall : GemmShape::kSquare;
  g.splits = small_splits(batch, g.K);
  g.kchunk = ceil_div(ceil_div(g.K, g.splits), 16) * 16;
  g.splits = ceil_div(g.K, g.kchunk);
  g.partial = partial;
  gemm_f64(g, st);
}

struct ReductionLayout {
  double *C, *X, *Bt, *part;
  size_t bytes;
};
ReductionLayout reduction_layout(char* base, int64_t N, int l) {
  const int64_t nl = N * l, h = N / 2 > 0 ? N / 2 : 1;
  Carve c{base};
  ReductionLayout L{};
  L.C = c.take(static_cast<size_t>(N) * nl * l);
  L.X = c.take(static_cast<size_t>(h) * nl * l);
  L.Bt = c.take(static_cast<size_t>(h) * l * l);
  size_t part = 0;  // the split count depends on the batch, so take the largest over the levels
  for (int64_t n = N; n > 1; n = (n + 1) / 2) {
    const size_t e = gemm_f64_partial_elems(l, l, small_splits(n / 2, nl), n / 2);
    if (e > part) part = e;
  }
  L.part = c.take(part);
  L.bytes = c.off;
  return L;
}

struct ExtrinsicLayout {
  double *G, *w, *V, *M, *X, *R, *wr, *vr, *part;
  size_t bytes;
};
ExtrinsicLayout extrinsic_layout(char* base, int64_t N, int l) {
  const int64_t nl = N * l;
  Carve c{base};
  ExtrinsicLayout L{};
  L.G = c.take(static_cast<size_t>(N) * l * l);
  L.w = c.take(static_cast<size_t>(N) * l);
  L.V = c.take(static_cast<size_t>(N) * l * l);
  L.M = c.take(static_cast<size_t>(nl) * l);
  L.X = c.take(static_cast<size_t>(nl) * l);
  L.R = c.take(static_cast<size_t>(l) * l);
  L.wr = c.take(l);
  L.vr = c.take(static_cast<size_t>(l) * l);
  const size_t pa = gemm_f64_partial_elems(l, l, small_splits(N, nl), N);
  const size_t pb = gemm_f64_partial_elems(l, l, small_splits(1, nl), 1);
  L.part = c.take(pa > pb ? pa : pb);
  L.bytes = c.off;
  return L;
}

}  // namespace

size_t combine_workspace_bytes(bool extrinsic, int64_t N, int l) {
  return extrinsic ? extrinsic_layout(nullptr, N, l).bytes : reduction_layout(nullptr, N, l).bytes;
}

void integrate_reduction(const double* Bs, int64_t N, int l, double* coef, char* ws,
                         size_t ws_bytes, hipStream_t st) {
  MCNLA_REQUIRE(l >= 1 && l <= kMaxL && N >= 1, MCNLA_ERR_INVALID, "bad reduction sizes");
  ReductionLayout L = reduction_layout(ws, N, l);
  MCNLA_REQUIRE(ws_bytes >= L.bytes, MCNLA_ERR_INVALID, "reduction workspace too small");
  const int64_t nl = N * l;
  hipLaunchKernelGGL(selectors_kernel, dim3(grid_for(N * nl * l)), dim3(kThreads), 0, st, L.C, N,
                     nl, l);
  MCNLA_LAUNCHED();
  for (int64_t n = N; n > 1; n = (n + 1) / 2) {
    const int64_t h = n / 2;
    // X_j = Bs C_{h+j} (Bs is symmetric: its rows serve as the M-contiguous operand)
    GemmF64Args x;
    x.M = nl;
    x.N = l;
    x.K = nl;
    x.A = Bs;
    x.lda = nl;
    x.a_mcontig = true;
    x.B = L.C + h * nl * l;
    x.ldb = l;
    x.C = L.X;
    x.ldc = l;
    x.shape = l <= 32 ? GemmShape::kTallN32 : GemmShape::kTallN64;
    x.batch = h;
    x.stride_a = 0;
    x.stride_b = nl * l;
    x.stride_c = nl * l;
    gemm_f64(x, st);
    // B_j = C_j' X_j
    GemmF64Args b;
    b.M = l;
    b.N = l;
    b.K = nl;
    b.A = L.C;
    b.lda = l;
    b.a_mcontig = true;
    b.B = L.X;
    b.ldb = l;
    b.C = L.Bt;
    b.ldc = l;
    b.stride_a = nl * l;
    b.stride_b = nl * l;
    b.stride_c = static_cast<int64_t>(l) * l;
    small_products(b, h, L.part, st);
    hipLaunchKernelGGL(reduction_pair_kernel, dim3(static_cast<unsigned>(h)), dim3(kThreads), 0, st,
                       L.Bt, L.C, nl, l, h);
    MCNLA_LAUNCHED();
  }
  MCNLA_HIP(hipMemcpyAsync(coef, L.C, static_cast<size_t>(nl) * l * sizeof(double),
                           hipMemcpyDeviceToDevice, st));
}

void integrate_extrinsic(const double* Bs, int64_t N, int l, double* coef, char* ws,
                         size_t ws_bytes, hipStream_t st) {
  MCNLA_REQUIRE(l >= 1 && l <= kMaxL && N >= 1, MCNLA_ERR_INVALID, "bad extrinsic-mean sizes");
  ExtrinsicLayout L = extrinsic_layout(ws, N, l);
  MCNLA_REQUIRE(ws_bytes >= L.bytes, MCNLA_ERR_INVALID, "extrinsic-mean workspace too small");
  const int64_t nl = N * l;
  // G_i = B_i B_i' with B_i the i-th block row of Bs (= block column by symmetry)
  GemmF64Args g;
  g.M = l;
  g.N = l;
  g.K = nl;
  g.A = Bs;
  g.lda = nl;
  g.a_mcontig = true;
  g.B = Bs;
  g.ldb = nl;
  g.b_ncontig = true;
  g.C = L.G;
  g.ldc = l;
  g.stride_a = l;
  g.stride_b = l;
  g.stride_c = static_cast<int64_t>(l) * l;
  small_products(g, N, L.part, st);
  batched_syev(L.G, L.w, L.V, N, l, st);
  hipLaunchKernelGGL(extrinsic_align_kernel, dim3(static_cast<unsigned>(N)), dim3(kThreads), 0, st,
                     Bs, L.V, nl, l, L.M);
  MCNLA_LAUNCHED();
  // Qs M has Gram R = M' Bs M; its left singular vectors are Qs M V_R diag(1/sqrt(w_R))
  GemmF64Args x;
  x.M = nl;
  x.N = l;
  x.K = nl;
  x.A = Bs;
  x.lda = nl;
  x.a_mcontig = true;
  x.B = L.M;
  x.ldb = l;
  x.C = L.X;
  x.ldc = l;
  x.shape = l <= 32 ? GemmShape::kTallN32 : GemmShape::kTallN64;
  gemm_f64(x, st);
  GemmF64Args r;
  r.M = l;
  r.N = l;
  r.K = nl;
  r.A = L.M;
  r.lda = l;
  r.a_mcontig = true;
  r.B = L.X;
  r.ldb = l;
  r.C = L.R;
  r.ldc = l;
  small_products(r, 1, L.part, st);
  batched_syev(L.R, L.wr, L.vr, 1, l, st);
  hipLaunchKernelGGL(rmul_invsqrt_kernel, dim3(grid_for(nl * l)), dim3(kThreads), 0, st, L.M, nl, l,
                     L.vr, L.wr, coef);
  MCNLA_LAUNCHED();
}

}  // namespace mcnla
