// Device-side building blocks for the small (l x l, l <= 64) FP64 problems of the iSVD path.
// Everything here runs inside ONE workgroup with its data in LDS.
#pragma once

#include "common.hpp"

namespace mcnla {

// Leading dimension used for an n x n matrix kept in LDS (odd stride avoids bank conflicts on
// column walks).
__host__ __device__ constexpr int lds_ld(int n) { return n + 1; }

// Parallel cyclic Jacobi eigen-decomposition of the symmetric n x n matrix `a` (LDS, ld = n+1).
// On return the diagonal of `a` holds the eigenvalues and the columns of `v` (LDS, ld = n+1) the
// eigenvectors.  If init_v is true `v` starts as the identity, otherwise `v` holds a basis W
// and `a` must already be W' A W (warm start; on return v = W * rotations).  Pairs follow the
// round-robin tournament ordering: n-1 rounds of n/2 disjoint rotations per sweep.  A pair is
// rotated only while |a_pq| > 1e-15 sqrt(|a_pp a_qq|) (the threshold of one-sided/two-sided
// Jacobi with high relative accuracy, Demmel & Veselic 1992), and the annihilated entry is set
// to zero; the iteration stops after a sweep without rotations.  `cs` needs 2*64 doubles, `pq`
// 2*64 ints, `red` 8 doubles of LDS.  All threads of the workgroup must call this.
__device__ inline int lds_jacobi(int n, double* a, double* v, bool init_v, double* cs, int* pq,
                                 double* red) {
  const int ld = lds_ld(n);
  const int tid = threadIdx.x, nt = blockDim.x;
  if (init_v)
    for (int e = tid; e < n * n; e += nt) v[(e / n) * ld + e % n] = (e / n == e % n) ? 1.0 : 0.0;
  const int np = (n + 1) & ~1;  // players, padded to even (index n is a bye when n is odd)
  const int npairs = np / 2;
  int* rotated = reinterpret_cast<int*>(red);
  int sweeps = 0;
  __syncthreads();
  for (int sweep = 0; sweep < 30; ++sweep) {
    if (tid == 0) *rotated = 0;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      // rotation parameters of this round's pairs
      for (int k = tid; k < npairs; k += nt) {
        const int x = (k == 0) ? 0 : ((k - 1 + round) % (np - 1)) + 1;
        const int yk = np - 1 - k;
        const int y = ((yk - 1 + round) % (np - 1)) + 1;
        int p = x < y ? x : y, q = x < y ? y : x;
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = a[p * ld + q];
          const double app = a[p * ld + p], aqq = a[q * ld + q];
          if (fabs(apq) > 1e-15 * sqrt(fabs(app * aqq)) && apq != 0.0) {
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            *rotated = 1;
          }
        } else {
          p = q = -1;
        }
        cs[2 * k] = c;
        cs[2 * k + 1] = s;
        pq[2 * k] = p;
        pq[2 * k + 1] = q;
      }
      __syncthreads();
      // rows p, q of a
      for (int e = tid; e < npairs * n; e += nt) {
        const int k = e / n, j = e % n;
        const int p = pq[2 * k], q = pq[2 * k + 1];
        if (p < 0 || cs[2 * k + 1] == 0.0) continue;
        const double c = cs[2 * k], s = cs[2 * k + 1];
        const double ap = a[p * ld + j], aq = a[q * ld + j];
        a[p * ld + j] = c * ap - s * aq;
        a[q * ld + j] = s * ap + c * aq;
      }
      __syncthreads();
      // columns p, q of a and of v
      for (int e = tid; e < npairs * n; e += nt) {
        const int k = e / n, i = e % n;
        const int p = pq[2 * k], q = pq[2 * k + 1];
        if (p < 0 || cs[2 * k + 1] == 0.0) continue;
        const double c = cs[2 * k], s = cs[2 * k + 1];
        const double ap = a[i * ld + p], aq = a[i * ld + q];
        a[i * ld + p] = c * ap - s * aq;
        a[i * ld + q] = s * ap + c * aq;
        const double vp = v[i * ld + p], vq = v[i * ld + q];
        v[i * ld + p] = c * vp - s * vq;
        v[i * ld + q] = s * vp + c * vq;
      }
      __syncthreads();
      // the rotation annihilates a_pq analytically; drop its rounding residue
      for (int k = tid; k < npairs; k += nt) {
        const int p = pq[2 * k], q = pq[2 * k + 1];
        if (p >= 0 && cs[2 * k + 1] != 0.0) {
          a[p * ld + q] = 0.0;
          a[q * ld + p] = 0.0;
        }
      }
      __syncthreads();
    }
    ++sweeps;
    const int any = *rotated;
    __syncthreads();
    if (!any) break;
  }
  return sweeps;
}

// c = a' * b for n x n LDS matrices (ld = n+1); c must not alias a or b.
__device__ inline void lds_matmul_tn(int n, const double* a, const double* b, double* c) {
  const int ld = lds_ld(n);
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    double s = 0.0;
    for (int q = 0; q < n; ++q) s += a[q * ld + i] * b[q * ld + j];
    c[i * ld + j] = s;
  }
}

// c = a * b for n x n LDS matrices (ld = n+1); c must not alias a or b.
__device__ inline void lds_matmul_nn(int n, const double* a, const double* b, double* c) {
  const int ld = lds_ld(n);
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    double s = 0.0;
    for (int q = 0; q < n; ++q) s += a[i * ld + q] * b[q * ld + j];
    c[i * ld + j] = s;
  }
}

// Sorts the eigenpairs left by lds_jacobi in DESCENDING order of eigenvalue and fixes each
// eigenvector's sign so its largest-magnitude entry is positive.  Writes w (n) and the sorted,
// sign-fixed vectors to vout (global, row-major n x n, vectors in columns).  `ord` needs n ints.
__device__ inline void lds_eig_sorted(int n, const double* a, const double* v, double* w,
                                      double* vout, int* ord) {
  const int ld = lds_ld(n);
  const int tid = threadIdx.x, nt = blockDim.x;
  // rank of each eigenvalue (ties broken by index) -> position in descending order; NaN ranks
  // last so that every rank is taken exactly once
  for (int i = tid; i < n; i += nt) {
    const double di = isnan(a[i * ld + i]) ? -INFINITY : a[i * ld + i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = isnan(a[j * ld + j]) ? -INFINITY : a[j * ld + j];
      if (dj > di || (dj == di && j < i)) ++r;
    }
    ord[r] = i;
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) {
    const int src = ord[j];
    w[j] = a[src * ld + src];
    int big = 0;
    for (int i = 1; i < n; ++i)
      if (fabs(v[i * ld + src]) > fabs(v[big * ld + src])) big = i;
    const double sg = v[big * ld + src] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) vout[i * n + j] = sg * v[i * ld + src];
  }
  __syncthreads();
}

}  // namespace mcnla
