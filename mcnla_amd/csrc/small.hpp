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
// to zero; the iteration stops after a sweep without rotations (a sweep that would rotate
// nothing is recognised by testing every pair once, and not run).  `cs` needs 2*64 doubles,
// `pq` 2*64 ints, `red` 8 doubles of LDS.  All threads of the workgroup must call this.
__device__ inline int lds_jacobi(int n, double* a, double* v, bool init_v, double* cs, int* pq,
                                 double* red) {
  const int ld = lds_ld(n);
  const int tid = threadIdx.x, nt = blockDim.x;
  if (init_v)
    for (int e = tid; e < n * n; e += nt) v[(e / n) * ld + e % n] = (e / n == e % n) ? 1.0 : 0.0;
  const int np = (n + 1) & ~1;  // players, padded to even (index n is a bye when n is odd)
  const int npairs = np / 2;
  int* rotated = reinterpret_cast<int*>(red);
  int* more = rotated + 1;
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
        }  // q == n: the bye of an odd n; p takes part in no rotation this round
        cs[2 * k] = c;
        cs[2 * k + 1] = s;
        pq[2 * k] = p;
        pq[2 * k + 1] = q;
      }
      __syncthreads();
      // A := J'AJ one 2 x 2 block (pair ka's rows, pair kb's columns) per work item: the block's
      // rows are rotated first, then its columns -- the same operations as rotating all rows and
      // then all columns, with no barrier in between since the blocks are disjoint.  The entry
      // a_pq that a rotation annihilates analytically is set to zero.  V := VJ in the same pass.
      for (int e = tid; e < npairs * npairs + npairs * n; e += nt) {
        if (e < npairs * npairs) {
          const int ka = e / npairs, kb = e % npairs;
          const int r0 = pq[2 * ka], r1 = pq[2 * ka + 1], c0 = pq[2 * kb], c1 = pq[2 * kb + 1];
          const double ca = cs[2 * ka], sa = cs[2 * ka + 1], cb = cs[2 * kb], sb = cs[2 * kb + 1];
          const bool ra = sa != 0.0, rb = sb != 0.0;  // a pair with the bye (index n) has s = 0
          if (!ra && !rb) continue;
          const bool v1 = r1 < n, w1 = c1 < n;  // r0, c0 < n always
          double x00 = a[r0 * ld + c0], x01 = w1 ? a[r0 * ld + c1] : 0.0;
          double x10 = v1 ? a[r1 * ld + c0] : 0.0, x11 = (v1 && w1) ? a[r1 * ld + c1] : 0.0;
          if (ra) {
            const double y00 = ca * x00 - sa * x10, y01 = ca * x01 - sa * x11;
            const double y10 = sa * x00 + ca * x10, y11 = sa * x01 + ca * x11;
            x00 = y00; x01 = y01; x10 = y10; x11 = y11;
          }
          if (rb) {
            const double y00 = cb * x00 - sb * x01, y01 = sb * x00 + cb * x01;
            const double y10 = cb * x10 - sb * x11, y11 = sb * x10 + cb * x11;
            x00 = y00; x01 = y01; x10 = y10; x11 = y11;
          }
          if (ka == kb && ra) {
            x01 = 0.0;
            x10 = 0.0;
          }
          a[r0 * ld + c0] = x00;
          if (w1) a[r0 * ld + c1] = x01;
          if (v1) a[r1 * ld + c0] = x10;
          if (v1 && w1) a[r1 * ld + c1] = x11;
        } else {
          const int f = e - npairs * npairs, k = f / n, i = f % n;
          const int p = pq[2 * k], q = pq[2 * k + 1];
          if (cs[2 * k + 1] == 0.0) continue;
          const double c = cs[2 * k], s = cs[2 * k + 1];
          const double vp = v[i * ld + p], vq = v[i * ld + q];
          v[i * ld + p] = c * vp - s * vq;
          v[i * ld + q] = s * vp + c * vq;
        }
      }
      __syncthreads();
    }
    ++sweeps;
    const int any = *rotated;
    __syncthreads();
    if (!any) break;
    // Would the next sweep rotate any pair?  (The same test as above, on the matrix it would
    // start from: if no pair passes it, that sweep changes nothing and ends the iteration, so it
    // is skipped -- the result is the same, bit for bit, one sweep sooner.)
    if (tid == 0) *more = 0;
    __syncthreads();
    bool pass = false;
    for (int e = tid; e < n * n; e += nt) {
      const int p = e / n, q = e % n;
      if (p >= q) continue;
      const double apq = a[p * ld + q];
      pass = pass ||
             (fabs(apq) > 1e-15 * sqrt(fabs(a[p * ld + p] * a[q * ld + q])) && apq != 0.0);
    }
    if (pass) *more = 1;
    __syncthreads();
    const int again = *more;
    if (!again) {
      ++sweeps;
      break;
    }
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

// ---- MFMA on LDS-resident square matrices ------------------------------------------------------
// np x np matrices (np a multiple of 16, at most 64) kept row-major in LDS with leading dimension
// ld.  Matrices of order l < np are embedded as diag(X, I) so that products, square roots and
// inverses leave the padding block equal to the identity.

__host__ __device__ constexpr int pad16(int n) { return (n + 15) & ~15; }

// One 16 x 16 tile of op(A) op(B) on v_mfma_f64_16x16x4f64 (the whole wave must call it).  Slot r
// of lane L holds element (i0 + L/16 + 4r, j0 + L%16) of the product.
template <bool TA, bool TB>
__device__ __forceinline__ d4 lds_mfma_tile(int np, int ld, const double* A, const double* B,
                                            int i0, int j0) {
  const int lane = threadIdx.x & (kWave - 1);
  const int r = lane & 15, kq = lane >> 4;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < np; k0 += 4) {
    const int k = k0 + kq;
    const double a = TA ? A[k * ld + i0 + r] : A[(i0 + r) * ld + k];
    const double b = TB ? B[(j0 + r) * ld + k] : B[k * ld + j0 + r];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  return acc;
}

// max that keeps a NaN once one was seen
__device__ __forceinline__ double nan_max(double a, double b) {
  if (a != a || b != b) return a + b;
  return a > b ? a : b;
}

// Coupled Newton-Schulz iteration for the principal square root (Higham, "Functions of
// Matrices", SIAM 2008, eq. (6.35)): Y0 = M, W0 = I, T = (3I - W Y)/2, Y <- Y T, W <- T W.  Y
// converges quadratically to M^{1/2} and W to M^{-1/2} when ||I - M|| < 1.  On entry Y holds the
// symmetric M; W is set to I here and T is scratch (all LDS, np x np, leading dimension ld).  The
// update is done in place: every wave keeps its output tiles in registers across a barrier.
// Stops once max|T - I| < 1e-8 (the following update leaves an error of about 3 (1e-8)^2).
// Returns the number of iterations, or -1 when that did not happen within maxit (M not positive
// definite, too ill-conditioned, or not finite).  `red` needs one double per wave.  Every thread
// of the workgroup (NW waves) must call this.
template <int NW>
__device__ inline int lds_ns_sqrt(int np, int ld, double* Y, double* W, double* T, double* red,
                                  int maxit) {
  constexpr int QMAX = (16 + NW - 1) / NW;  // output tiles per wave (at most 16 tiles, np <= 64)
  constexpr int nt = NW * kWave, nw = NW;
  const int tid = threadIdx.x;
  const int wave = tid / kWave, lane = tid & (kWave - 1);
  const int nt1 = np / 16, ntiles = nt1 * nt1;
  for (int e = tid; e < np * np; e += nt) W[(e / np) * ld + e % np] = (e / np == e % np) ? 1.0 : 0.0;
  __syncthreads();
  for (int it = 1; it <= maxit; ++it) {
    double dmax = 0.0;
    for (int tile = wave; tile < ntiles; tile += nw) {
      const int i0 = (tile / nt1) * 16, j0 = (tile % nt1) * 16;
      const d4 acc = lds_mfma_tile<false, false>(np, ld, W, Y, i0, j0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + (lane >> 4) + 4 * r, j = j0 + (lane & 15);
        const double id = (i == j) ? 1.0 : 0.0;
        const double x = 1.5 * id - 0.5 * acc[r];
        T[i * ld + j] = x;
        dmax = nan_max(dmax, fabs(x - id));
      }
    }
    for (int off = 32; off > 0; off >>= 1) dmax = nan_max(dmax, __shfl_xor(dmax, off, kWave));
    if (lane == 0) red[wave] = dmax;
    __syncthreads();
    double d = 0.0;
    for (int w = 0; w < nw; ++w) d = nan_max(d, red[w]);
    d4 ay[QMAX], aw[QMAX];
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int tile = wave + q * nw;
      if (tile < ntiles) {
        const int i0 = (tile / nt1) * 16, j0 = (tile % nt1) * 16;
        ay[q] = lds_mfma_tile<false, false>(np, ld, Y, T, i0, j0);
        aw[q] = lds_mfma_tile<false, false>(np, ld, T, W, i0, j0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int tile = wave + q * nw;
      if (tile < ntiles) {
        const int i0 = (tile / nt1) * 16, j0 = (tile % nt1) * 16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = i0 + (lane >> 4) + 4 * r, j = j0 + (lane & 15);
          Y[i * ld + j] = ay[q][r];
          W[i * ld + j] = aw[q][r];
        }
      }
    }
    __syncthreads();
    if (d < 1e-8) return it;
    if (!(d == d) || d > 1e6) return -1;
  }
  return -1;
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
