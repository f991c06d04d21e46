// C ABI of libmcnla_amd: stage handles (the reference's iSVD stage plugins) and the fused solver.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "solver.hpp"

namespace mcnla {

void omega_reference(uint32_t seed, bool col_block, int threads, int64_t len, double* out);
void column_indices_reference(uint32_t seed, bool col_block, int threads, int64_t len,
                              int64_t ncol, int64_t* out);

namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }
const char* get_error() { return g_error.c_str(); }

namespace {
thread_local Carve* g_carve = nullptr;
}
Carve* current_carve() { return g_carve; }
CarveScope::CarveScope(Carve* c) : prev_(g_carve) { g_carve = c; }
CarveScope::~CarveScope() { g_carve = prev_; }

namespace {

template <class F>
mcnla_status guard(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return MCNLA_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return MCNLA_ERR_INVALID;
  } catch (...) {
    set_error("unknown error");
    return MCNLA_ERR_INVALID;
  }
}

// ---- tuned GEMM plans: workgroup targets of the split-K FP64 GEMMs per solve shape ----
constexpr mcnla_tuning kDefaultTuning = {2304, 4096, 2304, 1024, 512, 512, 0};
// bits of mcnla_tuning::kernels
constexpr int64_t kKernGram = 1, kKernFormer = 2, kKernHOmega = 4, kKernOmegaP = 8,
                  kKernNbyL = 16, kKernQbar = 32, kKernAll = 63;
int kern_bit(const mcnla_tuning& t, int64_t bit) { return (t.kernels & bit) ? 1 : 0; }
constexpr int64_t kMaxTarget = 1 << 20;

struct TuneKey {
  int64_t v[6];  // route, precision, m, n, l, N
  bool operator<(const TuneKey& o) const {
    return std::lexicographical_compare(v, v + 6, o.v, o.v + 6);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, mcnla_tuning>& tune_table() {
  static std::map<TuneKey, mcnla_tuning> t;
  return t;
}
TuneKey tune_key(const mcnla_params& p, int route, int precision) {
  return TuneKey{{route, precision, p.nrow, p.ncol, p.rank + p.over_rank, p.num_sketch}};
}
bool valid_tuning(const mcnla_tuning& t) {
  for (int64_t v : {t.gram, t.former, t.h_omega, t.omega_p, t.n_by_l, t.qbar})
    if (v < 1 || v > kMaxTarget) return false;
  return t.kernels >= 0 && t.kernels <= kKernAll;
}
mcnla_tuning lookup_tuning(const TuneKey& k, bool* found) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = tune_table().find(k);
  if (found) *found = it != tune_table().end();
  return it == tune_table().end() ? kDefaultTuning : it->second;
}
void store_tuning(const TuneKey& k, const mcnla_tuning& t) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_table()[k] = t;
}

// Device allocation owned by a stage (grows, never shrinks).  While a CarveScope is active on the
// calling thread, ensure() takes its memory from that scope's region instead of hipMalloc (a
// solver created inside a server's arena); such a buffer is borrowed, is never freed, and may not
// grow once the scope has ended.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p && !borrowed) (void)hipFree(p);
  }
  void ensure(size_t n) {
    if (n <= bytes) return;
    if (Carve* c = current_carve()) {
      const size_t off = align_up(c->used, kCarveAlign);
      MCNLA_REQUIRE(c->counting || off + n <= c->cap, MCNLA_ERR_NOMEM,
                    "solver workspace exceeds its reserved region");
      // Counting scopes only size a layout: their pointers are never dereferenced.
      p = c->counting ? reinterpret_cast<void*>(kCarveAlign + off) : c->base + off;
      bytes = n;
      borrowed = true;
      c->used = off + n;
      return;
    }
    MCNLA_REQUIRE(!borrowed, MCNLA_ERR_STATE,
                  "a solver carved from a reserved region cannot grow its workspace");
    if (p) MCNLA_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    if (n == 0) return;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw Error{MCNLA_ERR_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed"};
    }
    bytes = n;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// hipEvent timing of a stage: time() and per-part times (stage_wrapper.hpp:170-192).
struct Timer {
  std::vector<hipEvent_t> ev;
  int used = 0;
  bool own = true;  // false: `ev` is a borrowed, fixed set (SolveRes::marks)
  ~Timer() {
    if (own)
      for (auto e : ev) (void)hipEventDestroy(e);
  }
  // Records into `n` events owned elsewhere from now on (at most n marks per run).
  void borrow(const hipEvent_t* evs, int n) {
    if (own)
      for (auto e : ev) (void)hipEventDestroy(e);
    ev.assign(evs, evs + n);
    own = false;
    used = 0;
  }
  void start(hipStream_t st) {
    used = 0;
    mark(st);
  }
  void mark(hipStream_t st) {
    if (used == static_cast<int>(ev.size())) {
      MCNLA_REQUIRE(own, MCNLA_ERR_STATE, "more stage marks than borrowed events");
      hipEvent_t e;
      MCNLA_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
    MCNLA_HIP(hipEventRecord(ev[used], st));
    ++used;
  }
  std::vector<double> parts() {
    std::vector<double> out;
    if (used < 2) return out;
    MCNLA_HIP(hipEventSynchronize(ev[used - 1]));
    for (int i = 1; i < used; ++i) {
      float ms = 0.f;
      MCNLA_HIP(hipEventElapsedTime(&ms, ev[i - 1], ev[i]));
      out.push_back(ms * 1e-3);
    }
    return out;
  }
};

void check_params(const mcnla_params* p) {
  MCNLA_REQUIRE(p != nullptr, MCNLA_ERR_INVALID, "params is null");
  MCNLA_REQUIRE(p->nrow >= 1 && p->ncol >= 1, MCNLA_ERR_INVALID, "matrix size must be positive");
  MCNLA_REQUIRE(p->rank >= 1 && p->over_rank >= 0 && p->num_sketch >= 1, MCNLA_ERR_INVALID,
                "need rank >= 1, over_rank >= 0, num_sketch >= 1");
  const int64_t l = p->rank + p->over_rank;
  MCNLA_REQUIRE(l <= p->nrow && l <= p->ncol, MCNLA_ERR_INVALID,
                "rank + over_rank must not exceed min(nrow, ncol)");
  MCNLA_REQUIRE(l <= kMaxL, MCNLA_ERR_INVALID, "rank + over_rank must be at most 64");
}

bool is_mixed(int precision) {
  return precision >= MCNLA_PREC_MIXED_BF16 && precision <= MCNLA_PREC_MIXED_F16X3;
}

bool valid_omega_mode(int mode) {
  return mode == MCNLA_OMEGA_PHILOX || mode == MCNLA_OMEGA_REFERENCE ||
         mode == MCNLA_OMEGA_PHILOX_F16;
}

// 16-bit format (0 BF16, 1 FP16) and product count of a mixed precision's sketch.
int sketch_fmt(int precision) {
  return precision >= MCNLA_PREC_MIXED_F16 ? 1 : 0;
}
int sketch_terms(int precision) {
  switch (precision) {
    case MCNLA_PREC_MIXED_BF16X3:
    case MCNLA_PREC_MIXED_F16X3: return 3;
    case MCNLA_PREC_MIXED_F16X2: return 2;
    default: return 1;
  }
}

void check_precision(int precision) {
  MCNLA_REQUIRE(precision == MCNLA_PREC_F64 || is_mixed(precision), MCNLA_ERR_INVALID,
                "unknown precision");
}

void check_a(const mcnla_params& p, const void* A, int64_t lda, int dtype, int order) {
  MCNLA_REQUIRE(A != nullptr, MCNLA_ERR_INVALID, "A is null");
  MCNLA_REQUIRE(dtype == MCNLA_F64 || dtype == MCNLA_F32, MCNLA_ERR_INVALID, "A dtype");
  MCNLA_REQUIRE(order == MCNLA_COL_MAJOR || order == MCNLA_ROW_MAJOR, MCNLA_ERR_INVALID, "A order");
  MCNLA_REQUIRE(lda >= (order == MCNLA_COL_MAJOR ? p.nrow : p.ncol), MCNLA_ERR_INVALID,
                "lda smaller than the leading dimension of A");
}

GemmShape tall_shape(int64_t ncols) {
  return ncols <= 32 ? GemmShape::kTallN32 : GemmShape::kTallN64;
}

// ------------------------------------------------------------------------------------------------
// Shared pipeline pieces.

// Y = A * Omega with Omega generated into `omega` (n x Nl FP64).
struct SketchCore {
  DevBuf omega, ws;
  DevBuf tail;  // split-K partials of the sketch's last rows (allocated on first use)
  std::vector<double> host;
  hipEvent_t gemm_beg = nullptr, gemm_end = nullptr;
  bool own_ev = true;     // false: the events belong to a SolveRes
  bool recorded = false;  // the events bracket a GEMM launched since the last reset
  ~SketchCore() {
    if (own_ev && gemm_beg) (void)hipEventDestroy(gemm_beg);
    if (own_ev && gemm_end) (void)hipEventDestroy(gemm_end);
  }
  static SketchBf16Args bf16_args(const mcnla_params& p, int precision) {
    SketchBf16Args s;
    s.m = p.nrow;
    s.n = p.ncol;
    s.nl = (p.rank + p.over_rank) * p.num_sketch;
    s.terms = sketch_terms(precision);
    s.fmt = sketch_fmt(precision);
    return s;
  }
  // Allocates Omega and, for a mixed precision, the BF16 panels of A and Omega.
  void prepare(const mcnla_params& p, int precision) {
    omega.ensure(p.ncol * (p.rank + p.over_rank) * p.num_sketch * sizeof(double));
    if (precision != MCNLA_PREC_F64) ws.ensure(sketch_bf16_workspace(bf16_args(p, precision)));
  }
  void generate(const mcnla_params& p, uint64_t seed, int mode, int threads, bool col_block,
                hipStream_t st) {
    const int64_t nl = (p.rank + p.over_rank) * p.num_sketch;
    const int64_t count = p.ncol * nl;
    omega.ensure(count * sizeof(double));
    if (mode == MCNLA_OMEGA_REFERENCE) {
      host.resize(count);
      omega_reference(static_cast<uint32_t>(seed), col_block, threads, count, host.data());
      MCNLA_HIP(hipMemcpyAsync(omega.p, host.data(), count * sizeof(double), hipMemcpyHostToDevice, st));
      MCNLA_HIP(hipStreamSynchronize(st));  // `host` is reused by the next call
    } else {
      omega_philox(omega.as<double>(), count, seed, st, mode == MCNLA_OMEGA_PHILOX_F16);
    }
  }
  // y_f64: a mixed-precision sketch writes its FP32-accumulated Y as FP64 (the fused solver's
  // Gram and basis products then read FP64); otherwise mixed Y is FP32.
  void project(const mcnla_params& p, int precision, const void* A, int64_t lda, int dtype,
               int order, void* Y, hipStream_t st, bool y_f64 = false) {
    const int64_t nl = (p.rank + p.over_rank) * p.num_sketch;
    if (!gemm_beg) {
      MCNLA_HIP(hipEventCreate(&gemm_beg));
      MCNLA_HIP(hipEventCreate(&gemm_end));
    }
    if (precision == MCNLA_PREC_F64) {
      GemmF64Args g;
      g.M = p.nrow;
      g.N = nl;
      g.K = p.ncol;
      g.A = A;
      g.lda = lda;
      g.a_mcontig = (order == MCNLA_COL_MAJOR);
      g.a_f32 = (dtype == MCNLA_F32);
      g.B = omega.p;
      g.ldb = nl;
      g.b_ncontig = true;
      g.C = static_cast<double*>(Y);
      g.ldc = nl;
      g.shape = GemmShape::kSquare;
      // A column-major FP64 A takes the DMA-tiled GEMM, one workgroup per CU: the rows of whole
      // rounds of the grid go as one product, the rest split K (their last round would otherwise
      // run a few full-length workgroups on a mostly idle device: 6256 tiles at 100000 x 10000,
      // Nl = 2048, are 24 rounds and 112 tiles).  Not for a solver carved from a reserved
      // region, which may not allocate the partials.
      int64_t ts = 1;
      const int64_t main = (g.a_mcontig && !g.a_f32 && !omega.borrowed)
                               ? gemm_f64_dma_tail(g.M, g.N, g.K, &ts)
                               : g.M;
      MCNLA_HIP(hipEventRecord(gemm_beg, st));
      if (ts > 1) {
        GemmF64Args t = g;
        t.M = p.nrow - main;
        t.A = static_cast<const double*>(A) + main;
        t.C = static_cast<double*>(Y) + main * nl;
        t.kchunk = ceil_div(ceil_div(g.K, ts), 16) * 16;
        t.splits = ceil_div(g.K, t.kchunk);
        tail.ensure(gemm_f64_partial_elems(t.M, t.N, t.splits) * sizeof(double));
        t.partial = tail.as<double>();
        g.M = main;
        gemm_f64(g, st);
        gemm_f64(t, st);
      } else {
        gemm_f64(g, st);
      }
      MCNLA_HIP(hipEventRecord(gemm_end, st));
      recorded = true;
    } else {
      SketchBf16Args s = bf16_args(p, precision);
      s.A = A;
      s.lda = lda;
      s.a_colmajor = (order == MCNLA_COL_MAJOR);
      s.a_f32 = (dtype == MCNLA_F32);
      s.omega = omega.as<double>();
      if (y_f64)
        s.Yd = static_cast<double*>(Y);
      else
        s.Y = static_cast<float*>(Y);
      s.ldy = nl;
      ws.ensure(sketch_bf16_workspace(s));
      // the rows past the grid's whole rounds split K (not in a solver carved from a reserved
      // region, which may not allocate the partials)
      if (!omega.borrowed) {
        const size_t tb = sketch_bf16_tail_bytes(s);
        if (tb) {
          tail.ensure(tb);
          s.tail = tail.as<float>();
          s.tail_bytes = tail.bytes;
        }
      }
      sketch_bf16(s, ws.p, ws.bytes, st, gemm_beg, gemm_end);
      recorded = true;
    }
  }
  double gemm_seconds() {
    if (!gemm_beg || !recorded) return 0.0;
    MCNLA_HIP(hipEventSynchronize(gemm_end));
    float ms = 0.f;
    MCNLA_HIP(hipEventElapsedTime(&ms, gemm_beg, gemm_end));
    return ms * 1e-3;
  }
};

// G = Y' * Y (Nl x Nl, FP64, full symmetric), Y row-major m x Nl.
struct GramCore {
  DevBuf partial;
  hipEvent_t beg = nullptr, end = nullptr;
  bool own_ev = true;     // false: the events belong to a SolveRes
  bool recorded = false;  // the events bracket a GEMM launched since the last reset
  ~GramCore() {
    if (own_ev && beg) (void)hipEventDestroy(beg);
    if (own_ev && end) (void)hipEventDestroy(end);
  }
  double gemm_seconds() {
    if (!beg || !recorded) return 0.0;
    MCNLA_HIP(hipEventSynchronize(end));
    float ms = 0.f;
    MCNLA_HIP(hipEventElapsedTime(&ms, beg, end));
    return ms * 1e-3;
  }
  // workgroups the split-K grid aims at (about 9 per CU by default), and the kernel (tile shape)
  int64_t target = kDefaultTuning.gram;
  int kern = 0;
  // Arguments of G = Y'Y except the pointers.
  GemmF64Args plan(bool y_f32, int64_t m, int64_t nl) const {
    // Lower-triangle tiles: 256 x 128 on the FP64 fast path (tm (tm+1) tiles), else 128 x 128.
    // Split K so that the grid fills whole rounds of about `target` workgroups.
    const int64_t tm = ceil_div(nl, 256), t = ceil_div(nl, 128);
    const int64_t tiles = (!y_f32 && nl >= 128 && kern == 0) ? tm * (tm + 1) : t * (t + 1) / 2;
    GemmF64Args g;
    g.M = nl;
    g.N = nl;
    g.K = m;
    g.lda = nl;
    g.a_mcontig = true;
    g.a_f32 = y_f32;
    g.ldb = nl;
    g.b_ncontig = true;
    g.b_f32 = y_f32;
    g.ldc = nl;
    g.lower_only = true;
    g.shape = GemmShape::kSquare;
    g.kernel = kern;
    g.splits = gemm_f64_splits_for(tiles, m, target);
    g.kchunk = ceil_div(m, g.splits);  // equal splits (the GEMM rounds them up where it must)
    g.splits = ceil_div(m, g.kchunk);
    return g;
  }
  void prepare(bool y_f32, int64_t m, int64_t nl) {
    const GemmF64Args g = plan(y_f32, m, nl);
    partial.ensure(gemm_f64_partial_elems(nl, nl, g.splits) * sizeof(double));
  }
  void run(const void* Y, bool y_f32, int64_t m, int64_t nl, double* G, hipStream_t st) {
    if (!beg) {
      MCNLA_HIP(hipEventCreate(&beg));
      MCNLA_HIP(hipEventCreate(&end));
    }
    GemmF64Args g = plan(y_f32, m, nl);
    g.ev_begin = beg;
    g.ev_end = end;
    g.A = Y;
    g.B = Y;
    g.C = G;
    prepare(y_f32, m, nl);
    g.partial = partial.as<double>();
    gemm_f64(g, st);
    recorded = true;
  }
};

// Z = A' Qbar (n x l), Wg = Z' Z, eig, and the outputs of either former.
struct FormCore {
  DevBuf Z, partial, Wg, eigw, eigv, Wk, S, U, V;
  // events around the Z = A'Qbar launch (the former's stream over A)
  hipEvent_t zbeg = nullptr, zend = nullptr;
  bool own_ev = true;     // false: the events belong to a SolveRes
  bool recorded = false;  // the events bracket a Z launch since the last reset
  FormCore() = default;
  FormCore(const FormCore&) = delete;
  FormCore& operator=(const FormCore&) = delete;
  ~FormCore() {
    if (own_ev && zbeg) (void)hipEventDestroy(zbeg);
    if (own_ev && zend) (void)hipEventDestroy(zend);
  }
  double z_seconds() {
    if (!zbeg || !recorded) return 0.0;
    MCNLA_HIP(hipEventSynchronize(zend));
    float ms = 0.f;
    MCNLA_HIP(hipEventElapsedTime(&ms, zbeg, zend));
    return ms * 1e-3;
  }
  // kind: MCNLA_FORM_*.  The symmetric former (square symmetric A) uses Wg = Z'Qbar = Qbar'A Qbar
  // in place of Z'Z and forms no V.
  // Split-K of Z = A'Qbar (K = m over ceil(n/128) row tiles) and of Wg = Z'Z (K = n).
  int64_t z_target = kDefaultTuning.former;  // workgroups of Z = A'Qbar's split-K grid
  int z_kern = 0;                            // and its kernel
  int64_t z_kchunk(int64_t m, int64_t n) const {
    return ceil_div(ceil_div(m, gemm_f64_splits_for(ceil_div(n, 128), m, z_target)), 16) * 16;
  }
  static int64_t w_kchunk(int64_t n) {
    return ceil_div(ceil_div(n, gemm_f64_splits_for(1, n, 64)), 16) * 16;
  }
  void prepare(const mcnla_params& p) {
    const int64_t m = p.nrow, n = p.ncol, k = p.rank, l = p.rank + p.over_rank;
    Z.ensure(n * l * sizeof(double));
    Wg.ensure(l * l * sizeof(double));
    eigw.ensure(l * sizeof(double));
    eigv.ensure(l * l * sizeof(double));
    Wk.ensure(l * k * sizeof(double));
    S.ensure(k * sizeof(double));
    U.ensure(m * k * sizeof(double));
    V.ensure(n * k * sizeof(double));
    partial.ensure(std::max(gemm_f64_partial_elems(n, l, ceil_div(m, z_kchunk(m, n))),
                            gemm_f64_partial_elems(l, l, ceil_div(n, w_kchunk(n)))) *
                   sizeof(double));
  }
  // z_ready: Z = A'Qbar is already in Z (the solver's Gram-of-A route forms it without A).
  // mixed: a mixed precision forms Z of an FP32 A on FP32 MFMA (FP32 sums within each split,
  // FP64 across splits); everything after Z stays FP64.
  void run(const mcnla_params& p, int kind, bool jobv, const void* A, int64_t lda, int dtype,
           int order, const double* Qbar, hipStream_t st, bool z_ready = false,
           bool mixed = false) {
    const bool svd = kind == MCNLA_FORM_SVD;
    const bool sym = kind == MCNLA_FORM_ROW_BLOCK_SYMMETRIC;
    if (sym) jobv = false;
    const int64_t m = p.nrow, n = p.ncol, k = p.rank;
    const int l = static_cast<int>(p.rank + p.over_rank);
    prepare(p);
    // Z = A' Qbar: element (j, r) of A' is A(r, j).
    GemmF64Args z;
    z.M = n;
    z.N = l;
    z.K = m;
    z.A = A;
    z.lda = lda;
    z.a_mcontig = (order == MCNLA_ROW_MAJOR);
    z.a_f32 = (dtype == MCNLA_F32);
    z.B = Qbar;
    z.ldb = l;
    z.C = Z.as<double>();
    z.ldc = l;
    z.shape = tall_shape(l);
    z.f32_mfma = mixed;
    z.kernel = z_kern;
    z.kchunk = z_kchunk(m, n);
    z.splits = ceil_div(m, z.kchunk);
    // Wg = Z' Z (symmetric former: Z' Qbar, Qbar having n = m rows)
    GemmF64Args w;
    w.M = l;
    w.N = l;
    w.K = n;
    w.A = Z.p;
    w.lda = l;
    w.a_mcontig = true;
    w.B = sym ? static_cast<const void*>(Qbar) : Z.p;
    w.ldb = l;
    w.C = Wg.as<double>();
    w.ldc = l;
    w.shape = GemmShape::kSmall;
    w.kchunk = w_kchunk(n);
    w.splits = ceil_div(n, w.kchunk);
    z.partial = partial.as<double>();
    w.partial = partial.as<double>();
    recorded = false;
    if (!z_ready) {
      if (!zbeg) {
        MCNLA_HIP(hipEventCreate(&zbeg));
        MCNLA_HIP(hipEventCreate(&zend));
      }
      z.ev_begin = zbeg;
      z.ev_end = zend;
      gemm_f64(z, st);
      recorded = true;
    }
    gemm_f64(w, st);
    former_small(Wg.as<double>(), l, static_cast<int>(k), eigw.as<double>(), eigv.as<double>(),
                 Wk.as<double>(), S.as<double>(), st, sym);
    // V = Z Wk (n x k); the SVD former takes sigma_j = ||Z w_j||.
    GemmF64Args v;
    v.M = n;
    v.N = k;
    v.K = l;
    v.A = Z.p;
    v.lda = l;
    v.a_mcontig = false;
    v.B = Wk.p;
    v.ldb = k;
    v.C = V.as<double>();
    v.ldc = k;
    v.shape = tall_shape(k);
    if (svd || jobv) gemm_f64(v, st);
    if (svd) column_norms(V.as<double>(), n, static_cast<int>(k), S.as<double>(), st);
    if (jobv) scale_columns(V.as<double>(), n, static_cast<int>(k), S.as<double>(), true, st);
    GemmF64Args u;
    if (svd) {
      // U column-major m x k == row-major k x m = Wk' Qbar'.
      u.M = k;
      u.N = m;
      u.K = l;
      u.A = Wk.p;
      u.lda = k;
      u.a_mcontig = true;
      u.B = Qbar;
      u.ldb = l;
      u.b_ncontig = false;
      u.C = U.as<double>();
      u.ldc = m;
      u.shape = (k <= 32) ? GemmShape::kWideM32 : GemmShape::kSquare;
    } else {
      u.M = m;
      u.N = k;
      u.K = l;
      u.A = Qbar;
      u.lda = l;
      u.a_mcontig = false;
      u.B = Wk.p;
      u.ldb = k;
      u.C = U.as<double>();
      u.ldc = k;
      u.shape = tall_shape(k);
    }
    gemm_f64(u, st);
  }
};

}  // namespace
}  // namespace mcnla

// ================================================================================================
// Stage objects.
struct mcnla_stage_s {
  enum Type { kSketcher, kOrthogonalizer, kIntegrator, kFormer } type;
  int kind = 0;
  int precision = 0;
  mcnla_params prm{};
  std::string name, part_names;
  mcnla::Timer timer;
  bool computed = false;
  explicit mcnla_stage_s(Type t) : type(t) {}
  virtual ~mcnla_stage_s() = default;
  int64_t l() const { return prm.rank + prm.over_rank; }
  int64_t nl() const { return l() * prm.num_sketch; }
};

namespace mcnla {
namespace {

struct Sketcher : mcnla_stage_s {
  uint64_t seed = 0;
  int mode = MCNLA_OMEGA_PHILOX;
  int threads = 1;
  int64_t exponent = 0;
  SketchCore core;
  DevBuf idx, partial;             // column-sampling indices; split-K partials of A'Y
  std::vector<int64_t> host_idx;
  Sketcher() : mcnla_stage_s(kSketcher) {}
  bool sampling() const {
    return kind == MCNLA_SKETCHER_COLUMN_SAMPLING || kind == MCNLA_SKETCHER_ROW_BLOCK_COLUMN_SAMPLING;
  }
};

// The row-block sketchers draw from Streams(seed); the others from Streams(seed, root, comm),
// whose per-rank seed_seq adds one step (streams.hpp:118-150).
bool col_block_chain(int kind) {
  return kind != MCNLA_SKETCHER_ROW_BLOCK_GAUSSIAN && kind != MCNLA_SKETCHER_ROW_BLOCK_COLUMN_SAMPLING;
}

struct Orthogonalizer : mcnla_stage_s {
  DevBuf W, w, V, T, partial, bad;
  Orthogonalizer() : mcnla_stage_s(kOrthogonalizer) {}
};

struct Integrator : mcnla_stage_s {
  int64_t maxiter = 256;
  double tol = 1e-4;
  int64_t iteration = 0;
  DevBuf Bs, ws, coef;
  GramCore gram;
  Integrator() : mcnla_stage_s(kIntegrator) {}
};

struct Former : mcnla_stage_s {
  bool jobv = true;
  FormCore core;
  Former() : mcnla_stage_s(kFormer) {}
};

const char* sketcher_name(int kind) {
  switch (kind) {
    case MCNLA_SKETCHER_COL_BLOCK_GAUSSIAN: return "Gaussian Projection Sketcher (Column-Block Version)";
    case MCNLA_SKETCHER_ROW_BLOCK_GAUSSIAN: return "Gaussian Projection Sketcher (Row-Block Version)";
    case MCNLA_SKETCHER_COLUMN_SAMPLING: return "Column Sampling Sketcher";
    case MCNLA_SKETCHER_ROW_BLOCK_COLUMN_SAMPLING: return "Column Sampling Sketcher (Row-Block Version)";
    default: return "Gaussian Projection Sketcher";
  }
}
const char* integrator_name(int kind) {
  switch (kind) {
    case MCNLA_INT_KN: return "Kolmogorov-Nagumo-Type Integrator";
    case MCNLA_INT_ROW_BLOCK_KN: return "Kolmogorov-Nagumo-Type Integrator (Row-Block Version)";
    case MCNLA_INT_ROW_BLOCK_GRAMIAN_KN:
      return "Kolmogorov-Nagumo-Type Integrator (Row-Block Gramian Version)";
    case MCNLA_INT_ROW_BLOCK_WY: return "Wen-Yin Line Search Integrator (Row-Block Version)";
    case MCNLA_INT_ROW_BLOCK_REDUCTION: return "Reduction Integrator (Row-Block Version)";
    case MCNLA_INT_ROW_BLOCK_EXTRINSIC_MEAN:
      return "Extrinsic Mean Integrator (Row-Block Version)";
    default: return "Wen-Yin Line Search Integrator (Row-Block Gramian Version)";
  }
}
const char* former_name(int kind) {
  switch (kind) {
    case MCNLA_FORM_SVD: return "SVD Former";
    case MCNLA_FORM_GRAMIAN: return "Gramian Former";
    case MCNLA_FORM_COL_BLOCK_GRAMIAN: return "Gramian Former (Column-Block Version)";
    case MCNLA_FORM_ROW_BLOCK_SYMMETRIC: return "Symmetric Former (Row-Block Version)";
    default: return "Gramian Former (Row-Block Version)";
  }
}

bool is_wen_yin(int kind) {
  return kind == MCNLA_INT_ROW_BLOCK_WY || kind == MCNLA_INT_ROW_BLOCK_GRAMIAN_WY;
}

bool is_combining(int kind) {
  return kind == MCNLA_INT_ROW_BLOCK_REDUCTION || kind == MCNLA_INT_ROW_BLOCK_EXTRINSIC_MEAN;
}

size_t integration_workspace(int kind, int64_t N, int l) {
  return is_combining(kind) ? combine_workspace_bytes(kind == MCNLA_INT_ROW_BLOCK_EXTRINSIC_MEAN, N, l)
                            : integrator_workspace_bytes(N, l);
}

// Runs integrator `kind` on the Gram matrix Bs (Nl x Nl) of an orthonormal collection; writes
// the Nl x l coefficients of Qbar.  Returns the iteration count (0 for the combining kinds).
int64_t run_integration(int kind, const double* Bs, int64_t N, int l, int64_t maxiter, double tol,
                        double* coef, DevBuf& ws, hipStream_t st, bool* conv, bool* ok,
                        IterationGraphs* graphs = nullptr) {
  ws.ensure(integration_workspace(kind, N, l));
  if (is_combining(kind)) {
    if (kind == MCNLA_INT_ROW_BLOCK_REDUCTION)
      integrate_reduction(Bs, N, l, coef, ws.as<char>(), ws.bytes, st);
    else
      integrate_extrinsic(Bs, N, l, coef, ws.as<char>(), ws.bytes, st);
    *conv = true;
    *ok = true;
    return 0;
  }
  GramIntegrator gi;
  gi.N = N;
  gi.l = l;
  gi.wen_yin = is_wen_yin(kind);
  gi.maxiter = maxiter;
  gi.tol = tol;
  gi.Bs = const_cast<double*>(Bs);
  gi.ws = ws.as<char>();
  gi.ws_bytes = ws.bytes;
  gi.graphs = graphs;
  return integrate_gram(gi, coef, st, conv, ok);
}

hipStream_t S(void* stream) { return static_cast<hipStream_t>(stream); }

// Orthogonalize the collection in place: Q_i = Y_i V_i S_i^{-1/2}.
int orthogonalize_inplace(Orthogonalizer& o, void* Qs, hipStream_t st) {
  const auto& p = o.prm;
  const int l = static_cast<int>(o.l());
  const int64_t N = p.num_sketch, m = p.nrow, nl = o.nl();
  const bool f32 = is_mixed(o.precision);
  o.W.ensure(N * l * l * sizeof(double));
  o.w.ensure(N * l * sizeof(double));
  o.V.ensure(N * l * l * sizeof(double));
  o.T.ensure(N * l * l * sizeof(double));
  o.bad.ensure(sizeof(int));
  GemmF64Args g;
  g.M = l;
  g.N = l;
  g.K = m;
  g.A = Qs;
  g.lda = nl;
  g.a_mcontig = true;
  g.a_f32 = f32;
  g.B = Qs;
  g.ldb = nl;
  g.b_f32 = f32;
  g.C = o.W.as<double>();
  g.ldc = l;
  g.shape = (l <= 32) ? GemmShape::kSmall : GemmShape::kSquare;
  g.batch = N;
  g.stride_a = l;
  g.stride_b = l;
  g.stride_c = static_cast<int64_t>(l) * l;
  g.splits = gemm_f64_splits_for(N, m, 2048);
  g.kchunk = ceil_div(ceil_div(m, g.splits), 16) * 16;
  g.splits = ceil_div(m, g.kchunk);
  o.partial.ensure(gemm_f64_partial_elems(l, l, g.splits, N) * sizeof(double));
  g.partial = o.partial.as<double>();
  gemm_f64(g, st);
  MCNLA_HIP(hipMemsetAsync(o.bad.p, 0, sizeof(int), st));
  batched_syev(o.W.as<double>(), o.w.as<double>(), o.V.as<double>(), N, l, st);
  orth_transforms(o.w.as<double>(), o.V.as<double>(), o.T.as<double>(), N, l, o.bad.as<int>(), st);
  apply_blockdiag(Qs, f32, m, N, l, o.T.as<double>(), st);
  int bad = 0;
  MCNLA_HIP(hipMemcpyAsync(&bad, o.bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  MCNLA_HIP(hipStreamSynchronize(st));
  return bad;
}

// Runs a Gram-based integrator over an orthonormal collection Qs; Qbar = Qs * coef.
mcnla_status integrate(Integrator& it, const void* Qs, double* Qbar, hipStream_t st) {
  const auto& p = it.prm;
  const int l = static_cast<int>(it.l());
  const int64_t N = p.num_sketch, m = p.nrow, nl = it.nl();
  const bool f32 = is_mixed(it.precision);
  it.Bs.ensure(nl * nl * sizeof(double));
  it.coef.ensure(nl * l * sizeof(double));
  it.timer.start(st);
  it.gram.run(Qs, f32, m, nl, it.Bs.as<double>(), st);
  it.timer.mark(st);
  bool conv = false, ok = true;
  it.iteration = run_integration(it.kind, it.Bs.as<double>(), N, l, it.maxiter, it.tol,
                                 it.coef.as<double>(), it.ws, st, &conv, &ok);
  it.timer.mark(st);
  GemmF64Args q;
  q.M = m;
  q.N = l;
  q.K = nl;
  q.A = Qs;
  q.lda = nl;
  q.a_mcontig = false;
  q.a_f32 = f32;
  q.B = it.coef.p;
  q.ldb = l;
  q.C = Qbar;
  q.ldc = l;
  q.shape = tall_shape(l);
  gemm_f64(q, st);
  it.timer.mark(st);
  it.computed = true;
  if (!ok) {
    set_error("non-finite value in the integrator's l x l step");
    return MCNLA_ERR_NUMERIC;
  }
  if (!conv && it.maxiter > 0) {
    set_error("integrator reached maxiter without meeting the tolerance");
    return MCNLA_ERR_NOT_CONVERGED;
  }
  return MCNLA_OK;
}

}  // namespace
}  // namespace mcnla

// ================================================================================================
// Solver.
struct mcnla_solver_s {
  mcnla_params prm{};
  mcnla_solver_config cfg{};
  mcnla::SketchCore sketch;
  mcnla::GramCore gram;
  mcnla::FormCore form;
  mcnla::DevBuf Y, G, W, w, V, T, Bs, tmp, ws, coef, M2, Qbar, bad;
  mcnla::DevBuf L0, T0, W0;  // first block's Cholesky factor, eigen transform, and rotation
  // mcnla_solver_solve_host: the device copy of a host A, the copy stream and one event per chunk
  mcnla::DevBuf Acopy;
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> chunk_ev;
  // Gram-of-A route: A as row-major FP64, H = A'A, P = H Omega, X = Omega (T M), split-K scratch
  mcnla::DevBuf A64, H, P, X, gp;
  int route = MCNLA_ROUTE_SKETCH;
  mcnla_tuning tune{};  // split-K workgroup targets of the FP64 GEMMs
  mcnla::Timer timer;
  // HIP objects of a solve (the side stream of the first block's eigen-decomposition, which runs
  // beside the integrator, and the timing events): the solver's own set, created on its first
  // solve, or a set lent by the caller of solver_create_in
  mcnla::SolveRes own_res;
  mcnla::SolveRes* res = &own_res;
  std::unique_ptr<mcnla::IterationGraphs> graphs;  // cfg.graph: recorded iteration chunks
  int64_t iteration = 0;
  bool computed = false;
  bool converged = false;
  ~mcnla_solver_s() {
    // the cores' and timer's events are the SolveRes's; drop them before the set goes
    timer.ev.clear();
    own_res.destroy();
    for (hipEvent_t e : chunk_ev) (void)hipEventDestroy(e);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
  }
  // Creates the solver's own set if it has none, and points the timer and cores at it.
  void bind_res() {
    if (!res->ready()) res->create();
    timer.borrow(res->marks, mcnla::SolveRes::kMarks);
    sketch.gemm_beg = res->sketch_beg;
    sketch.gemm_end = res->sketch_end;
    sketch.own_ev = false;
    sketch.recorded = false;
    gram.beg = res->gram_beg;
    gram.end = res->gram_end;
    gram.own_ev = false;
    gram.recorded = false;
    form.zbeg = res->former_beg;
    form.zend = res->former_end;
    form.own_ev = false;
    form.recorded = false;
  }
};

namespace mcnla {

namespace {
// The route a solve of this shape takes: the one with less FP64-MFMA-equivalent work.  The sketch
// route costs the Gram m Nl^2 plus the sketch 2 m n Nl (at about 1/16 of that on BF16 MFMA for
// the mixed precisions); the Gram-of-A route costs A'A (m n^2), H Omega (2 n^2 Nl) and
// Omega'(H Omega) (n Nl^2).
int resolve_route(const mcnla_params& p, const mcnla_solver_config& c) {
  if (c.route == MCNLA_ROUTE_SKETCH || c.route == MCNLA_ROUTE_GRAM_OF_A) return c.route;
  const double m = static_cast<double>(p.nrow), n = static_cast<double>(p.ncol);
  const double nl = static_cast<double>((p.rank + p.over_rank) * p.num_sketch);
  const double sketch = 2.0 * m * n * nl *
                        (c.precision == MCNLA_PREC_F64 ? 1.0 : sketch_terms(c.precision) / 16.0);
  const double via_y = m * nl * nl + sketch;
  const double via_a = m * n * n + 2.0 * n * n * nl + n * nl * nl;
  return via_a < via_y ? MCNLA_ROUTE_GRAM_OF_A : MCNLA_ROUTE_SKETCH;
}

// P = H Omega (n x Nl): H symmetric, so its rows serve as the M-contiguous operand.
GemmF64Args plan_h_omega(int64_t n, int64_t nl, int64_t target, int kern) {
  GemmF64Args g;
  g.kernel = kern;
  g.M = n;
  g.N = nl;
  g.K = n;
  g.lda = n;
  g.a_mcontig = true;
  g.ldb = nl;
  g.b_ncontig = true;
  g.ldc = nl;
  g.shape = GemmShape::kSquare;
  g.splits = gemm_f64_splits_for(kern == 0 ? ceil_div(n, 256) * ceil_div(nl, 128)
                                           : ceil_div(n, 128) * ceil_div(nl, 128),
                                 n, target);
  g.kchunk = ceil_div(ceil_div(n, g.splits), 16) * 16;
  g.splits = ceil_div(n, g.kchunk);
  return g;
}

// Split-K factor of the n x l products X = Omega (T M) (K = Nl) and Z = H X (K = n): n is small
// on this route, so without splits they would run on n/128 workgroups.
int64_t n_by_l_splits(int64_t n, int64_t K, int64_t target) {
  return gemm_f64_splits_for(ceil_div(n, 128), K, target);
}

// Split-K factor of Qbar = Y (T M) (m x l, K = Nl) on the sketch route: m/128 row tiles.
int64_t qbar_splits(int64_t m, int64_t nl, int64_t target) {
  return gemm_f64_splits_for(ceil_div(m, 128), nl, target);
}

// G = Omega' P (Nl x Nl, symmetric): lower-triangle tiles.
GemmF64Args plan_omega_p(int64_t n, int64_t nl, int64_t target, int kern) {
  GemmF64Args g;
  g.kernel = kern;
  g.M = nl;
  g.N = nl;
  g.K = n;
  g.lda = nl;
  g.a_mcontig = true;
  g.ldb = nl;
  g.b_ncontig = true;
  g.ldc = nl;
  g.lower_only = true;
  g.shape = GemmShape::kSquare;
  const int64_t tm = ceil_div(nl, 256), t = ceil_div(nl, 128);
  // by default about 4 rounds of the grid without splitting K where the tiles allow it: a split
  // writes and re-reads Nl^2 doubles of partials (1.6 GB per 3 splits at Nl = 8192)
  g.splits = gemm_f64_splits_for(nl >= 128 && kern == 0 ? tm * (tm + 1) : t * (t + 1) / 2, n,
                                 target);
  g.kchunk = ceil_div(ceil_div(n, g.splits), 16) * 16;
  g.splits = ceil_div(n, g.kchunk);
  return g;
}

// Sets the solver's GEMM plan and sizes the split-K scratch it needs (grows only).
void apply_tuning(mcnla_solver_s* s, const mcnla_tuning& t) {
  s->tune = t;
  s->gram.target = t.gram;
  s->gram.kern = (t.kernels & kKernGram) ? 1 : 0;
  s->form.z_target = t.former;
  s->form.z_kern = (t.kernels & kKernFormer) ? 1 : 0;
  const int64_t l = s->prm.rank + s->prm.over_rank, nl = l * s->prm.num_sketch;
  const int64_t m = s->prm.nrow, n = s->prm.ncol;
  const bool via_a = s->route == MCNLA_ROUTE_GRAM_OF_A;
  if (via_a)
    s->gp.ensure(std::max({gemm_f64_partial_elems(n, nl, plan_h_omega(n, nl, t.h_omega, kern_bit(t, kKernHOmega)).splits),
                           gemm_f64_partial_elems(nl, nl, plan_omega_p(n, nl, t.omega_p, kern_bit(t, kKernOmegaP)).splits),
                           gemm_f64_partial_elems(n, l, n_by_l_splits(n, nl, t.n_by_l)),
                           gemm_f64_partial_elems(n, l, n_by_l_splits(n, n, t.n_by_l))}) *
                 sizeof(double));
  s->gram.prepare(false, m, via_a ? n : nl);
  if (!via_a)  // the Qbar product's split-K partials reuse the Gram's (free by then)
    s->gram.partial.ensure(gemm_f64_partial_elems(m, l, qbar_splits(m, nl, t.qbar)) *
                           sizeof(double));
  s->form.prepare(s->prm);
}

// Validates the configuration and allocates every device buffer the solve uses (the stages'
// initialize()), so that a solve itself allocates nothing.
std::unique_ptr<mcnla_solver_s> new_solver(const mcnla_params* params,
                                           const mcnla_solver_config* cfg) {
  check_params(params);
  MCNLA_REQUIRE(cfg, MCNLA_ERR_INVALID, "null argument");
  check_precision(cfg->precision);
  MCNLA_REQUIRE(cfg->sketcher >= MCNLA_SKETCHER_COL_BLOCK_GAUSSIAN &&
                cfg->sketcher <= MCNLA_SKETCHER_GAUSSIAN, MCNLA_ERR_INVALID,
                "the solver sketches with a Gaussian projection (column sampling is a stage "
                "sketcher)");
  MCNLA_REQUIRE(cfg->integrator >= MCNLA_INT_KN && cfg->integrator <= MCNLA_INT_ROW_BLOCK_EXTRINSIC_MEAN,
                MCNLA_ERR_INVALID, "unknown integrator kind");
  MCNLA_REQUIRE(cfg->former >= MCNLA_FORM_SVD && cfg->former <= MCNLA_FORM_ROW_BLOCK_SYMMETRIC,
                MCNLA_ERR_INVALID, "unknown former kind");
  MCNLA_REQUIRE(cfg->former != MCNLA_FORM_ROW_BLOCK_SYMMETRIC || params->nrow == params->ncol,
                MCNLA_ERR_INVALID, "the symmetric former needs a square A");
  MCNLA_REQUIRE(valid_omega_mode(cfg->omega_mode), MCNLA_ERR_INVALID, "unknown omega mode");
  MCNLA_REQUIRE(cfg->maxiter >= 0 && cfg->tol >= 0.0, MCNLA_ERR_INVALID, "bad maxiter/tol");
  auto s = std::make_unique<mcnla_solver_s>();
  s->prm = *params;
  s->cfg = *cfg;
  const int64_t l = params->rank + params->over_rank, N = params->num_sketch, nl = l * N;
  const int64_t m = params->nrow;
  s->route = resolve_route(*params, *cfg);
  const bool via_a = s->route == MCNLA_ROUTE_GRAM_OF_A;
  const int64_t n = params->ncol;
  if (!via_a) {
    // Y is FP64 in every precision: the mixed sketch accumulates in FP32 and stores FP64, so that
    // the FP64 Gram and basis products stream it at full rate.
    s->Y.ensure(m * nl * sizeof(double));
  } else {
    s->A64.ensure(m * n * sizeof(double));
    s->H.ensure(n * n * sizeof(double));
    s->P.ensure(n * nl * sizeof(double));
    s->X.ensure(n * l * sizeof(double));
  }
  apply_tuning(s.get(), lookup_tuning(tune_key(*params, s->route, cfg->precision), nullptr));
  s->G.ensure(nl * nl * sizeof(double));
  s->Bs.ensure(nl * nl * sizeof(double));
  s->tmp.ensure(nl * l * sizeof(double));
  s->W.ensure(N * l * l * sizeof(double));
  s->w.ensure(N * l * sizeof(double));
  s->V.ensure(N * l * l * sizeof(double));
  s->T.ensure(N * l * l * sizeof(double));
  s->ws.ensure(integration_workspace(cfg->integrator, N, static_cast<int>(l)));
  s->coef.ensure(nl * l * sizeof(double));
  s->M2.ensure(nl * l * sizeof(double));
  s->Qbar.ensure(m * l * sizeof(double));
  s->bad.ensure(sizeof(int));
  s->L0.ensure(l * l * sizeof(double));
  s->T0.ensure(l * l * sizeof(double));
  s->W0.ensure(l * l * sizeof(double));
  s->sketch.prepare(*params, via_a ? static_cast<int>(MCNLA_PREC_F64) : cfg->precision);
  return s;
}
}  // namespace

size_t solver_required_bytes(const mcnla_params* params, const mcnla_solver_config* cfg) {
  Carve c;
  c.counting = true;
  CarveScope scope(&c);
  new_solver(params, cfg);
  return align_up(c.used, kCarveAlign);
}

mcnla_solver solver_create_in(const mcnla_params* params, const mcnla_solver_config* cfg,
                              void* base, size_t bytes, SolveRes* res) {
  MCNLA_REQUIRE(base != nullptr && reinterpret_cast<uintptr_t>(base) % kCarveAlign == 0,
                MCNLA_ERR_INVALID, "region base must be 256-byte aligned");
  MCNLA_REQUIRE(res == nullptr || res->ready(), MCNLA_ERR_STATE, "lent SolveRes is not created");
  Carve c;
  c.base = static_cast<char*>(base);
  c.cap = bytes;
  CarveScope scope(&c);
  auto s = new_solver(params, cfg);
  if (res) s->res = res;
  return s.release();
}

void SolveRes::create() {
  if (ready()) return;
  try {
    MCNLA_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    MCNLA_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    MCNLA_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    for (auto& e : marks) MCNLA_HIP(hipEventCreate(&e));
    for (hipEvent_t* e : {&sketch_beg, &sketch_end, &gram_beg, &gram_end, &former_beg, &former_end})
      MCNLA_HIP(hipEventCreate(e));
  } catch (...) {
    destroy();
    throw;
  }
}

void SolveRes::destroy() {
  for (hipEvent_t* e :
       {&fork, &join, &sketch_beg, &sketch_end, &gram_beg, &gram_end, &former_beg, &former_end})
    if (*e) (void)hipEventDestroy(*e), *e = nullptr;
  for (auto& e : marks)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  if (side) (void)hipStreamDestroy(side), side = nullptr;
}

}  // namespace mcnla

using namespace mcnla;

namespace mcnla {

void solve_check(mcnla_solver s, const void* A, int64_t lda, int a_dtype, int a_order) {
  MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "solver is null");
  check_a(s->prm, A, lda, a_dtype, a_order);
}

// Stage 1 of a fused solve on `st` (after bind_res and the timer's start): Omega, then the
// sketch Y = A Omega, or H = A'A on the Gram-of-A route.
void solve_sketch(mcnla_solver_s* s, const void* A, int64_t lda, int a_dtype, int a_order,
                  hipStream_t st, SolveCtx& cx) {
  const auto& p = s->prm;
  const auto& c = s->cfg;
  const int64_t m = p.nrow;
  cx.ldar = p.ncol;
  cx.rotate = !is_combining(c.integrator);
  const bool via_a = s->route == MCNLA_ROUTE_GRAM_OF_A;
  const int64_t n = p.ncol;
  // 1. sketch: Y = A * Omega; on the Gram-of-A route, H = A'A (A as row-major FP64) instead
  s->sketch.generate(p, c.seed, c.omega_mode, c.omega_threads,
                     col_block_chain(c.sketcher), st);
  if (!via_a) {
    s->sketch.project(p, c.precision, A, lda, a_dtype, a_order, s->Y.p, st, true);
  } else {
    if (a_dtype == MCNLA_F64 && a_order == MCNLA_ROW_MAJOR && lda == n) {
      cx.Ar = static_cast<const double*>(A);
    } else {
      to_rowmajor_f64(A, lda, a_dtype == MCNLA_F32, a_order == MCNLA_COL_MAJOR, m, n,
                      s->A64.as<double>(), st);
      cx.Ar = s->A64.as<double>();
    }
    s->gram.run(cx.Ar, false, m, n, s->H.as<double>(), st);
  }
}

// Stage 2: G = Y'Y (or Omega'(H Omega)), the per-block transforms T and Bs = T'GT.
void solve_orthogonalize(mcnla_solver_s* s, hipStream_t st, SolveCtx& cx) {
  const auto& p = s->prm;
  const auto& c = s->cfg;
  const int l = static_cast<int>(p.rank + p.over_rank);
  const int64_t N = p.num_sketch, nl = l * N, m = p.nrow, n = p.ncol;
  const bool via_a = s->route == MCNLA_ROUTE_GRAM_OF_A;
  (void)c;
  // 2. orthogonalize through the Gram matrix: G = Y'Y, T_i from G_ii, Bs = T' G T.  The
  // reference's T_i = V_i S_i^{-1/2} (eig of G_ii) needs a Jacobi eigen-solve per block; for
  // the KN and WY integrators, which are equivariant in the basis of the first sketch (Qbar
  // turns into Qbar R when Q_0 turns into Q_0 R) and depend on the other sketches only through
  // their spans, T_i = L_i^{-T} (Cholesky, G_ii = L_i L_i') gives the same subspaces, and the
  // reference's basis is restored at the end by W_0 = L_0' V_0 S_0^{-1/2} = Q_0' Q_0^ref, with
  // the one eigen-solve it needs running on a second stream beside the integrator.  The
  // combining integrators depend on every basis and keep the eigen transforms.
  if (!via_a) {
    s->gram.run(s->Y.p, false, m, nl, s->G.as<double>(), st);
  } else {  // G = Omega' (H Omega)
    GemmF64Args hp = plan_h_omega(n, nl, s->tune.h_omega, kern_bit(s->tune, kKernHOmega));
    hp.A = s->H.p;
    hp.B = s->sketch.omega.p;
    hp.C = s->P.as<double>();
    hp.partial = s->gp.as<double>();
    gemm_f64(hp, st);
    GemmF64Args pg = plan_omega_p(n, nl, s->tune.omega_p, kern_bit(s->tune, kKernOmegaP));
    pg.A = s->sketch.omega.p;
    pg.B = s->P.p;
    pg.C = s->G.as<double>();
    pg.partial = s->gp.as<double>();
    gemm_f64(pg, st);
  }
  gram_diag_blocks(s->G.as<double>(), nl, s->W.as<double>(), N, l, st);
  MCNLA_HIP(hipMemsetAsync(s->bad.p, 0, sizeof(int), st));
  if (cx.rotate) {
    hipStream_t side = s->res->side;
    MCNLA_HIP(hipEventRecord(s->res->fork, st));
    MCNLA_HIP(hipStreamWaitEvent(side, s->res->fork, 0));
    batched_syev(s->W.as<double>(), s->w.as<double>(), s->V.as<double>(), 1, l, side);
    orth_transforms(s->w.as<double>(), s->V.as<double>(), s->T0.as<double>(), 1, l,
                    s->bad.as<int>(), side);
    MCNLA_HIP(hipEventRecord(s->res->join, side));
    cholesky_transforms(s->W.as<double>(), N, l, s->T.as<double>(), s->L0.as<double>(),
                        s->bad.as<int>(), st);
  } else {
    batched_syev(s->W.as<double>(), s->w.as<double>(), s->V.as<double>(), N, l, st);
    orth_transforms(s->w.as<double>(), s->V.as<double>(), s->T.as<double>(), N, l,
                    s->bad.as<int>(), st);
  }
  congruence_blockdiag(s->G.as<double>(), s->T.as<double>(), s->Bs.as<double>(), N, l, st);
}

void solve_begin(mcnla_solver_s* s, const void* A, int64_t lda, int a_dtype, int a_order,
                 hipStream_t st, SolveCtx& cx, bool timed) {
  s->bind_res();
  if (timed) s->timer.start(st);
  solve_sketch(s, A, lda, a_dtype, a_order, st, cx);
  if (timed) s->timer.mark(st);
  solve_orthogonalize(s, st, cx);
  if (timed) s->timer.mark(st);
}

void solve_finish(mcnla_solver_s* s, const void* A, int64_t lda, int a_dtype, int a_order,
                  hipStream_t st, const SolveCtx& cx, bool timed) {
  const auto& p = s->prm;
  const auto& c = s->cfg;
  const int l = static_cast<int>(p.rank + p.over_rank);
  const int64_t N = p.num_sketch, nl = l * N, m = p.nrow, n = p.ncol;
  const bool via_a = s->route == MCNLA_ROUTE_GRAM_OF_A;
  (void)N;
  const double* cf = s->coef.as<double>();
  if (cx.rotate) {
    MCNLA_HIP(hipStreamWaitEvent(st, s->res->join, 0));
    GemmF64Args wr;  // W_0 = L_0' T_0^ref
    wr.M = l;
    wr.N = l;
    wr.K = l;
    wr.A = s->L0.p;
    wr.lda = l;
    wr.a_mcontig = true;
    wr.B = s->T0.p;
    wr.ldb = l;
    wr.C = s->W0.as<double>();
    wr.ldc = l;
    wr.shape = GemmShape::kSmall;
    gemm_f64(wr, st);
    GemmF64Args cr;  // coef W_0 (Nl x l) into the scratch
    cr.M = nl;
    cr.N = l;
    cr.K = l;
    cr.A = s->coef.p;
    cr.lda = l;
    cr.a_mcontig = false;
    cr.B = s->W0.p;
    cr.ldb = l;
    cr.C = s->tmp.as<double>();
    cr.ldc = l;
    cr.shape = tall_shape(l);
    gemm_f64(cr, st);
    cf = s->tmp.as<double>();
  }
  GemmF64Args tm;  // M2_i = T_i * coef_i  (batched l x l)
  tm.M = l;
  tm.N = l;
  tm.K = l;
  tm.A = s->T.p;
  tm.lda = l;
  tm.a_mcontig = false;
  tm.B = cf;
  tm.ldb = l;
  tm.C = s->M2.as<double>();
  tm.ldc = l;
  tm.shape = GemmShape::kSmall;
  tm.batch = N;
  tm.stride_a = static_cast<int64_t>(l) * l;
  tm.stride_b = static_cast<int64_t>(l) * l;
  tm.stride_c = static_cast<int64_t>(l) * l;
  gemm_f64(tm, st);
  GemmF64Args q;
  q.M = m;
  q.N = l;
  q.K = nl;
  q.A = s->Y.p;
  q.lda = nl;
  q.a_mcontig = false;
  q.a_f32 = false;
  q.B = s->M2.p;
  q.ldb = l;
  q.C = s->Qbar.as<double>();
  q.ldc = l;
  q.shape = tall_shape(l);
  if (!via_a) {
    q.kernel = kern_bit(s->tune, kKernQbar);
    q.splits = qbar_splits(m, nl, s->tune.qbar);
    if (q.splits > 1) {
      q.kchunk = ceil_div(ceil_div(nl, q.splits), 16) * 16;
      q.partial = s->gram.partial.as<double>();
    }
  }
  if (via_a) {  // X = Omega (T M) (n x l), then Qbar = A X
    GemmF64Args x = q;
    x.M = n;
    x.A = s->sketch.omega.p;
    x.C = s->X.as<double>();
    x.splits = n_by_l_splits(n, nl, s->tune.n_by_l);
    x.kernel = kern_bit(s->tune, kKernNbyL);
    x.partial = s->gp.as<double>();
    gemm_f64(x, st);
    q.K = n;
    q.B = s->X.p;
    if (a_order == MCNLA_ROW_MAJOR) {  // stream the caller's A (FP32 reads half the bytes)
      q.A = A;
      q.a_f32 = a_dtype == MCNLA_F32;
      // a mixed precision forms this basis of an FP32 A on FP32 MFMA (sigma come from Z = H X,
      // which stays FP64)
      q.f32_mfma = c.precision != MCNLA_PREC_F64;
      q.lda = lda;
    } else {
      q.A = cx.Ar;
      q.lda = cx.ldar;
    }
  }
  gemm_f64(q, st);
  if (timed) s->timer.mark(st);
  // 4. form; on the Gram-of-A route Z = A'Qbar = H X without another pass over A
  if (via_a) {
    GemmF64Args z;
    z.M = n;
    z.N = l;
    z.K = n;
    z.A = s->H.p;
    z.lda = n;
    z.a_mcontig = false;
    z.B = s->X.p;
    z.ldb = l;
    z.C = s->form.Z.as<double>();
    z.ldc = l;
    z.shape = tall_shape(l);
    z.splits = n_by_l_splits(n, n, s->tune.n_by_l);
    z.kernel = kern_bit(s->tune, kKernNbyL);
    z.partial = s->gp.as<double>();
    gemm_f64(z, st);
  }
  s->form.run(p, c.former, c.jobv != 0, A, lda, a_dtype, a_order, s->Qbar.as<double>(), st,
              via_a, c.precision != MCNLA_PREC_F64);
  if (timed) s->timer.mark(st);
  s->computed = true;
}

void read_bad(mcnla_solver_s* s, int* bad, hipStream_t st) {
  MCNLA_HIP(hipMemcpyAsync(bad, s->bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
}

mcnla_status solve_status(const mcnla_solver_s* s, const SolveCtx& cx, int bad) {
  if (bad || !cx.ok) {
    set_error(bad ? "a sketch Gram matrix is not positive definite"
                  : "non-finite value in the integrator's l x l step");
    return MCNLA_ERR_NUMERIC;
  }
  if (!cx.conv && s->cfg.maxiter > 0) {
    set_error("integrator reached maxiter without meeting the tolerance");
    return MCNLA_ERR_NOT_CONVERGED;
  }
  return MCNLA_OK;
}

namespace {
bool is_kn(int kind) { return !is_combining(kind) && !is_wen_yin(kind); }
static_assert(MCNLA_MAX_BATCH <= kMaxKnGroup, "a batch's KN members form one group");
}  // namespace

bool solver_is_kn(mcnla_solver s) { return is_kn(s->cfg.integrator); }

GramIntegrator solver_kn_integrator(mcnla_solver s) {
  const auto& p = s->prm;
  const auto& c = s->cfg;
  const int l = static_cast<int>(p.rank + p.over_rank);
  s->ws.ensure(integration_workspace(c.integrator, p.num_sketch, l));
  GramIntegrator gi;
  gi.N = p.num_sketch;
  gi.l = l;
  gi.maxiter = c.maxiter;
  gi.tol = c.tol;
  gi.Bs = s->Bs.as<double>();
  gi.ws = s->ws.as<char>();
  gi.ws_bytes = s->ws.bytes;
  return gi;
}

double* solver_coef(mcnla_solver s) { return s->coef.as<double>(); }

void solver_set_integration(mcnla_solver s, int64_t iterations, bool converged) {
  s->iteration = iterations;
  s->converged = converged;
}

}  // namespace mcnla

namespace mcnla {

void solve_batch(int count, const mcnla_solver* solvers, const void* const* A, const int64_t* lda,
                 const int* a_dtype, const int* a_order, mcnla_status* status,
                 std::vector<std::string>* errors, hipStream_t st) {
  MCNLA_REQUIRE(count >= 1 && count <= MCNLA_MAX_BATCH, MCNLA_ERR_INVALID,
                "a batch holds 1 to 16 solves");
  MCNLA_REQUIRE(solvers && A && lda && a_dtype && a_order && status, MCNLA_ERR_INVALID,
                "null argument");
  for (int i = 0; i < count; ++i) {
    MCNLA_REQUIRE(solvers[i], MCNLA_ERR_INVALID, "solver is null");
    for (int j = 0; j < i; ++j)
      MCNLA_REQUIRE(solvers[j] != solvers[i], MCNLA_ERR_INVALID,
                    "a solver appears twice in the batch");
    check_a(solvers[i]->prm, A[i], lda[i], a_dtype[i], a_order[i]);
  }
  std::vector<SolveCtx> cx(count);
  for (int i = 0; i < count; ++i)
    solve_begin(solvers[i], A[i], lda[i], a_dtype[i], a_order[i], st, cx[i], false);
  // KN members in one group; the others one after another
  std::vector<GramIntegrator> group;
  std::vector<double*> coefs;
  std::vector<int> who;
  for (int i = 0; i < count; ++i) {
    mcnla_solver_s* s = solvers[i];
    const auto& p = s->prm;
    const auto& c = s->cfg;
    const int l = static_cast<int>(p.rank + p.over_rank);
    if (is_kn(c.integrator)) {
      group.push_back(solver_kn_integrator(s));
      coefs.push_back(s->coef.as<double>());
      who.push_back(i);
    } else {
      s->iteration = run_integration(c.integrator, s->Bs.as<double>(), p.num_sketch, l,
                                     c.maxiter, c.tol, s->coef.as<double>(), s->ws, st,
                                     &cx[i].conv, &cx[i].ok);
      s->converged = cx[i].conv;
    }
  }
  if (!group.empty()) {
    const int g = static_cast<int>(group.size());
    std::vector<int64_t> iters(g);
    std::unique_ptr<bool[]> conv(new bool[g]), ok(new bool[g]);
    integrate_gram_group(group.data(), coefs.data(), g, st, iters.data(), conv.get(), ok.get());
    for (int k = 0; k < g; ++k) {
      mcnla_solver_s* s = solvers[who[k]];
      s->iteration = iters[k];
      s->converged = conv[k];
      cx[who[k]].conv = conv[k];
      cx[who[k]].ok = ok[k];
    }
  }
  for (int i = 0; i < count; ++i)
    solve_finish(solvers[i], A[i], lda[i], a_dtype[i], a_order[i], st, cx[i], false);
  std::vector<int> bad(count, 0);
  for (int i = 0; i < count; ++i) read_bad(solvers[i], &bad[i], st);
  MCNLA_HIP(hipStreamSynchronize(st));
  if (errors) errors->assign(count, std::string());
  std::string first_error;
  for (int i = 0; i < count; ++i) {
    // the members' GEMM timing events may have been re-recorded by later members
    solvers[i]->sketch.recorded = false;
    solvers[i]->gram.recorded = false;
    status[i] = solve_status(solvers[i], cx[i], bad[i]);
    if (status[i] != MCNLA_OK) {
      if (errors) (*errors)[i] = get_error();
      if (first_error.empty()) first_error = get_error();
    }
  }
  set_error(first_error);
}

}  // namespace mcnla

extern "C" {

const char* mcnla_last_error(void) { return get_error(); }
const char* mcnla_version(void) { return "mcnla_amd 0.1.0 (gfx950)"; }

mcnla_status mcnla_device_count(int* count) {
  return guard([&] {
    MCNLA_REQUIRE(count, MCNLA_ERR_INVALID, "count is null");
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess || *count == 0) {
      *count = 0;
      (void)hipGetLastError();
      throw Error{MCNLA_ERR_NO_DEVICE, "no HIP device"};
    }
    return MCNLA_OK;
  });
}

mcnla_status mcnla_set_device(int device) {
  return guard([&] {
    MCNLA_HIP(hipSetDevice(device));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_malloc(void** ptr, size_t bytes) {
  return guard([&] {
    MCNLA_REQUIRE(ptr, MCNLA_ERR_INVALID, "ptr is null");
    *ptr = nullptr;
    if (bytes == 0) return MCNLA_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      *ptr = nullptr;
      throw Error{MCNLA_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
    }
    return MCNLA_OK;
  });
}

mcnla_status mcnla_free(void* ptr) {
  return guard([&] {
    if (ptr) MCNLA_HIP(hipFree(ptr));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_memcpy_htod(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    if (bytes) MCNLA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    if (bytes) MCNLA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  return guard([&] {
    if (bytes) MCNLA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_memset(void* ptr, int value, size_t bytes) {
  return guard([&] {
    if (bytes) MCNLA_HIP(hipMemset(ptr, value, bytes));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_stream_create(void** stream) {
  return guard([&] {
    hipStream_t s;
    MCNLA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_stream_destroy(void* stream) {
  return guard([&] {
    if (stream) MCNLA_HIP(hipStreamDestroy(S(stream)));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_stream_synchronize(void* stream) {
  return guard([&] {
    MCNLA_HIP(hipStreamSynchronize(S(stream)));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_event_create(void** event) {
  return guard([&] {
    MCNLA_REQUIRE(event, MCNLA_ERR_INVALID, "event is null");
    hipEvent_t e;
    MCNLA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    *event = e;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_event_destroy(void* event) {
  return guard([&] {
    if (event) MCNLA_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_event_record(void* event, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(event, MCNLA_ERR_INVALID, "event is null");
    MCNLA_HIP(hipEventRecord(static_cast<hipEvent_t>(event), S(stream)));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_event_synchronize(void* event) {
  return guard([&] {
    MCNLA_REQUIRE(event, MCNLA_ERR_INVALID, "event is null");
    MCNLA_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_host_malloc(void** ptr, size_t bytes) {
  return guard([&] {
    MCNLA_REQUIRE(ptr, MCNLA_ERR_INVALID, "ptr is null");
    *ptr = nullptr;
    if (bytes == 0) return MCNLA_OK;
    if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      *ptr = nullptr;
      throw Error{MCNLA_ERR_NOMEM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed"};
    }
    return MCNLA_OK;
  });
}

mcnla_status mcnla_host_free(void* ptr) {
  return guard([&] {
    if (ptr) MCNLA_HIP(hipHostFree(ptr));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_memcpy_htod_async(void* dst, const void* src, size_t bytes, void* stream) {
  return guard([&] {
    if (bytes) MCNLA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_device_synchronize(void) {
  return guard([&] {
    MCNLA_HIP(hipDeviceSynchronize());
    return MCNLA_OK;
  });
}

mcnla_status mcnla_mem_info(size_t* free_bytes, size_t* total_bytes) {
  return guard([&] {
    MCNLA_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return MCNLA_OK;
  });
}

// ---------------------------------------------------------------------------------------------
mcnla_status mcnla_sketcher_create(const mcnla_params* params, int kind, int precision,
                                   mcnla_stage* out) {
  return guard([&] {
    check_params(params);
    check_precision(precision);
    MCNLA_REQUIRE(out, MCNLA_ERR_INVALID, "out is null");
    MCNLA_REQUIRE(kind >= MCNLA_SKETCHER_COL_BLOCK_GAUSSIAN &&
                  kind <= MCNLA_SKETCHER_ROW_BLOCK_COLUMN_SAMPLING,
                  MCNLA_ERR_INVALID, "unknown sketcher kind");
    auto s = std::make_unique<Sketcher>();
    s->kind = kind;
    s->precision = precision;
    s->prm = *params;
    s->name = sketcher_name(kind);
    s->part_names = "random generating / projection";
    if (s->sampling()) {
      MCNLA_REQUIRE(precision == MCNLA_PREC_F64, MCNLA_ERR_INVALID,
                    "column sampling copies A: use MCNLA_PREC_F64");
      s->idx.ensure(s->nl() * sizeof(int64_t));
    } else {
      s->core.omega.ensure(params->ncol * s->nl() * sizeof(double));
    }
    *out = s.release();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_sketcher_set_seed(mcnla_stage s, uint64_t seed, int omega_mode, int threads) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kSketcher, MCNLA_ERR_INVALID, "not a sketcher");
    MCNLA_REQUIRE(valid_omega_mode(omega_mode), MCNLA_ERR_INVALID, "unknown omega mode");
    auto* k = static_cast<Sketcher*>(s);
    k->seed = seed;
    k->mode = omega_mode;
    k->threads = threads < 1 ? 1 : threads;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_sketcher_run(mcnla_stage s, const void* A, int64_t lda, int a_dtype,
                                int a_order, void* Y, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kSketcher, MCNLA_ERR_INVALID, "not a sketcher");
    auto* k = static_cast<Sketcher*>(s);
    check_a(k->prm, A, lda, a_dtype, a_order);
    MCNLA_REQUIRE(Y, MCNLA_ERR_INVALID, "Y is null");
    hipStream_t st = S(stream);
    const auto& p = k->prm;
    const int64_t nl = k->nl();
    k->timer.start(st);
    if (k->sampling()) {  // column_sampling_sketcher.hpp:84-96
      k->host_idx.resize(nl);
      column_indices_reference(static_cast<uint32_t>(k->seed), col_block_chain(k->kind), k->threads,
                               nl, p.ncol, k->host_idx.data());
      MCNLA_HIP(hipMemcpyAsync(k->idx.p, k->host_idx.data(), nl * sizeof(int64_t),
                               hipMemcpyHostToDevice, st));
      MCNLA_HIP(hipStreamSynchronize(st));  // host_idx is reused by the next run
      k->timer.mark(st);
      gather_columns(A, lda, a_dtype == MCNLA_F32, a_order == MCNLA_COL_MAJOR, p.nrow,
                     k->idx.as<int64_t>(), nl, static_cast<double*>(Y), st);
      k->timer.mark(st);
      k->computed = true;
      return MCNLA_OK;
    }
    k->core.generate(p, k->seed, k->mode, k->threads, col_block_chain(k->kind), st);
    k->timer.mark(st);
    k->core.project(p, k->precision, A, lda, a_dtype, a_order, Y, st);
    for (int64_t q = 0; q < k->exponent; ++q) {  // gaussian_projection_sketcher.hpp:96-100
      GemmF64Args t;  // Omega := A' Y  (n x Nl)
      t.M = p.ncol;
      t.N = nl;
      t.K = p.nrow;
      t.A = A;
      t.lda = lda;
      t.a_mcontig = (a_order == MCNLA_ROW_MAJOR);  // row j of A' is column j of A
      t.a_f32 = (a_dtype == MCNLA_F32);
      t.B = Y;
      t.ldb = nl;
      t.b_ncontig = true;
      t.C = k->core.omega.as<double>();
      t.ldc = nl;
      t.shape = GemmShape::kSquare;
      t.splits = gemm_f64_splits_for(ceil_div(p.ncol, 128) * ceil_div(nl, 128), p.nrow, 1024);
      if (t.splits > 1) {
        k->partial.ensure(gemm_f64_partial_elems(p.ncol, nl, t.splits) * sizeof(double));
        t.partial = k->partial.as<double>();
      }
      gemm_f64(t, st);
      k->core.project(p, k->precision, A, lda, a_dtype, a_order, Y, st);  // Y := A Omega
    }
    k->timer.mark(st);
    k->computed = true;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_sketcher_set_exponent(mcnla_stage s, int64_t exponent) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kSketcher, MCNLA_ERR_INVALID, "not a sketcher");
    auto* k = static_cast<Sketcher*>(s);
    MCNLA_REQUIRE(exponent >= 0, MCNLA_ERR_INVALID, "negative exponent");
    MCNLA_REQUIRE(exponent == 0 || k->kind == MCNLA_SKETCHER_GAUSSIAN, MCNLA_ERR_INVALID,
                  "only the Gaussian projection sketcher has power iterations");
    MCNLA_REQUIRE(exponent == 0 || k->precision == MCNLA_PREC_F64, MCNLA_ERR_INVALID,
                  "power iterations need MCNLA_PREC_F64");
    k->exponent = exponent;
    k->computed = false;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_sketcher_omega(mcnla_stage s, const double** omega) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kSketcher, MCNLA_ERR_INVALID, "not a sketcher");
    MCNLA_REQUIRE(s->computed, MCNLA_ERR_STATE, "sketcher has not run");
    MCNLA_REQUIRE(!static_cast<Sketcher*>(s)->sampling(), MCNLA_ERR_INVALID,
                  "a column-sampling sketcher has no Omega");
    *omega = static_cast<Sketcher*>(s)->core.omega.as<double>();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_sketcher_indices(mcnla_stage s, const int64_t** idx) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kSketcher, MCNLA_ERR_INVALID, "not a sketcher");
    MCNLA_REQUIRE(s->computed, MCNLA_ERR_STATE, "sketcher has not run");
    MCNLA_REQUIRE(static_cast<Sketcher*>(s)->sampling(), MCNLA_ERR_INVALID,
                  "not a column-sampling sketcher");
    MCNLA_REQUIRE(idx, MCNLA_ERR_INVALID, "idx is null");
    *idx = static_cast<Sketcher*>(s)->idx.as<int64_t>();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_column_indices_reference(uint64_t seed, int sketcher_kind, int threads,
                                            int64_t len, int64_t ncol, int64_t* out) {
  return guard([&] {
    MCNLA_REQUIRE(out || len == 0, MCNLA_ERR_INVALID, "out is null");
    MCNLA_REQUIRE(len >= 0 && ncol >= 1, MCNLA_ERR_INVALID, "bad length or column count");
    MCNLA_REQUIRE(threads >= 1 && threads <= 1024, MCNLA_ERR_INVALID, "threads out of range");
    column_indices_reference(static_cast<uint32_t>(seed), col_block_chain(sketcher_kind), threads,
                             len, ncol, out);
    return MCNLA_OK;
  });
}

mcnla_status mcnla_omega_reference(uint64_t seed, int sketcher_kind, int threads, int64_t len,
                                   double* out) {
  return guard([&] {
    MCNLA_REQUIRE(out || len == 0, MCNLA_ERR_INVALID, "out is null");
    MCNLA_REQUIRE(len >= 0, MCNLA_ERR_INVALID, "negative length");
    MCNLA_REQUIRE(threads >= 1 && threads <= 1024, MCNLA_ERR_INVALID, "threads out of range");
    omega_reference(static_cast<uint32_t>(seed), col_block_chain(sketcher_kind), threads, len, out);
    return MCNLA_OK;
  });
}

mcnla_status mcnla_orthogonalizer_create(const mcnla_params* params, int kind, int precision,
                                         mcnla_stage* out) {
  return guard([&] {
    check_params(params);
    check_precision(precision);
    MCNLA_REQUIRE(out, MCNLA_ERR_INVALID, "out is null");
    MCNLA_REQUIRE(kind == MCNLA_ORTH_GRAMIAN || kind == MCNLA_ORTH_ROW_BLOCK_GRAMIAN,
                  MCNLA_ERR_INVALID, "unknown orthogonalizer kind");
    auto o = std::make_unique<Orthogonalizer>();
    o->kind = kind;
    o->precision = precision;
    o->prm = *params;
    o->name = kind == MCNLA_ORTH_GRAMIAN ? "Gramian Orthogonalizer"
                                         : "Gramian Orthogonalizer (Row-Block Version)";
    o->part_names = "computing";
    *out = o.release();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_orthogonalizer_run(mcnla_stage s, void* Qs, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kOrthogonalizer, MCNLA_ERR_INVALID,
                  "not an orthogonalizer");
    MCNLA_REQUIRE(Qs, MCNLA_ERR_INVALID, "Qs is null");
    auto* o = static_cast<Orthogonalizer*>(s);
    hipStream_t st = S(stream);
    o->timer.start(st);
    const int bad = orthogonalize_inplace(*o, Qs, st);
    o->timer.mark(st);
    o->computed = true;
    if (bad) {
      set_error(std::to_string(bad) + " sketch Gram matrices are not positive definite");
      return MCNLA_ERR_NUMERIC;
    }
    return MCNLA_OK;
  });
}

mcnla_status mcnla_integrator_create(const mcnla_params* params, int kind, int precision,
                                     int64_t maxiter, double tol, mcnla_stage* out) {
  return guard([&] {
    check_params(params);
    check_precision(precision);
    MCNLA_REQUIRE(out, MCNLA_ERR_INVALID, "out is null");
    MCNLA_REQUIRE(kind >= MCNLA_INT_KN && kind <= MCNLA_INT_ROW_BLOCK_EXTRINSIC_MEAN, MCNLA_ERR_INVALID,
                  "unknown integrator kind");
    MCNLA_REQUIRE(maxiter >= 0 && maxiter < (1LL << 30), MCNLA_ERR_INVALID, "bad maxiter");
    MCNLA_REQUIRE(tol >= 0.0, MCNLA_ERR_INVALID, "tolerance must be non-negative");
    auto it = std::make_unique<Integrator>();
    it->kind = kind;
    it->precision = precision;
    it->prm = *params;
    it->maxiter = maxiter;
    it->tol = tol;
    it->name = integrator_name(kind);
    it->part_names = "initializing / iterating / forming";
    *out = it.release();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_integrator_run(mcnla_stage s, const void* Qs, double* Qbar, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kIntegrator, MCNLA_ERR_INVALID,
                  "not an integrator");
    MCNLA_REQUIRE(Qs && Qbar, MCNLA_ERR_INVALID, "null buffer");
    return integrate(*static_cast<Integrator*>(s), Qs, Qbar, S(stream));
  });
}

mcnla_status mcnla_integrator_iteration(mcnla_stage s, int64_t* iteration) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kIntegrator, MCNLA_ERR_INVALID,
                  "not an integrator");
    *iteration = static_cast<Integrator*>(s)->iteration;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_former_create(const mcnla_params* params, int kind, int jobv, int precision,
                                 mcnla_stage* out) {
  return guard([&] {
    check_params(params);
    check_precision(precision);
    MCNLA_REQUIRE(out, MCNLA_ERR_INVALID, "out is null");
    MCNLA_REQUIRE(kind >= MCNLA_FORM_SVD && kind <= MCNLA_FORM_ROW_BLOCK_SYMMETRIC, MCNLA_ERR_INVALID,
                  "unknown former kind");
    auto f = std::make_unique<Former>();
    f->kind = kind;
    f->precision = precision;
    f->prm = *params;
    f->jobv = jobv != 0;
    f->name = former_name(kind);
    f->part_names = "projection / decomposition / forming";
    *out = f.release();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_former_run(mcnla_stage s, const void* A, int64_t lda, int a_dtype,
                              int a_order, const double* Qbar, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kFormer, MCNLA_ERR_INVALID, "not a former");
    auto* f = static_cast<Former*>(s);
    check_a(f->prm, A, lda, a_dtype, a_order);
    MCNLA_REQUIRE(Qbar, MCNLA_ERR_INVALID, "Qbar is null");
    MCNLA_REQUIRE(f->kind != MCNLA_FORM_ROW_BLOCK_SYMMETRIC || f->prm.nrow == f->prm.ncol,
                  MCNLA_ERR_INVALID, "the symmetric former needs a square A");
    hipStream_t st = S(stream);
    f->timer.start(st);
    f->core.run(f->prm, f->kind, f->jobv, A, lda, a_dtype, a_order, Qbar, st, false,
                f->precision != MCNLA_PREC_F64);
    f->timer.mark(st);
    f->computed = true;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_former_results(mcnla_stage s, const double** Sv, const double** U,
                                  const double** V) {
  return guard([&] {
    MCNLA_REQUIRE(s && s->type == mcnla_stage_s::kFormer, MCNLA_ERR_INVALID, "not a former");
    auto* f = static_cast<Former*>(s);
    MCNLA_REQUIRE(f->computed, MCNLA_ERR_STATE, "former has not run");
    if (Sv) *Sv = f->core.S.as<double>();
    if (U) *U = f->core.U.as<double>();
    if (V)
      *V = (f->jobv && f->kind != MCNLA_FORM_ROW_BLOCK_SYMMETRIC) ? f->core.V.as<double>() : nullptr;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_stage_destroy(mcnla_stage s) {
  return guard([&] {
    delete s;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_stage_time(mcnla_stage s, double* total, double* parts, int max_parts,
                              int* nparts) {
  return guard([&] {
    MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "stage is null");
    MCNLA_REQUIRE(s->computed, MCNLA_ERR_STATE, "stage has not run");
    const auto v = s->timer.parts();
    double t = 0.0;
    for (size_t i = 0; i < v.size(); ++i) {
      t += v[i];
      if (parts && static_cast<int>(i) < max_parts) parts[i] = v[i];
    }
    if (total) *total = t;
    if (nparts) *nparts = static_cast<int>(v.size());
    return MCNLA_OK;
  });
}

const char* mcnla_stage_name(mcnla_stage s) { return s ? s->name.c_str() : ""; }
const char* mcnla_stage_part_names(mcnla_stage s) { return s ? s->part_names.c_str() : ""; }

// ---------------------------------------------------------------------------------------------
void mcnla_solver_config_default(mcnla_solver_config* cfg) {
  if (!cfg) return;
  cfg->sketcher = MCNLA_SKETCHER_COL_BLOCK_GAUSSIAN;
  cfg->orthogonalizer = MCNLA_ORTH_ROW_BLOCK_GRAMIAN;
  cfg->integrator = MCNLA_INT_ROW_BLOCK_KN;
  cfg->former = MCNLA_FORM_SVD;
  cfg->precision = MCNLA_PREC_F64;
  cfg->jobv = 1;
  cfg->maxiter = 256;
  cfg->tol = 1e-4;
  cfg->seed = 0;
  cfg->omega_mode = MCNLA_OMEGA_PHILOX;
  cfg->omega_threads = 1;
  cfg->route = MCNLA_ROUTE_AUTO;
  cfg->graph = 0;
}

mcnla_status mcnla_solver_create(const mcnla_params* params, const mcnla_solver_config* cfg,
                                 mcnla_solver* out) {
  return guard([&] {
    MCNLA_REQUIRE(out, MCNLA_ERR_INVALID, "null argument");
    *out = new_solver(params, cfg).release();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_required_bytes(const mcnla_params* params, const mcnla_solver_config* cfg,
                                         size_t* bytes) {
  return guard([&] {
    MCNLA_REQUIRE(bytes, MCNLA_ERR_INVALID, "null argument");
    *bytes = solver_required_bytes(params, cfg);
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_destroy(mcnla_solver s) {
  return guard([&] {
    delete s;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_workspace_bytes(mcnla_solver s, size_t* bytes) {
  return guard([&] {
    MCNLA_REQUIRE(s && bytes, MCNLA_ERR_INVALID, "null argument");
    size_t b = 0;
    for (const DevBuf* d : {&s->Y, &s->A64, &s->H, &s->P, &s->X, &s->gp, &s->L0, &s->T0, &s->W0,
                            &s->G, &s->W, &s->w, &s->V, &s->T, &s->Bs, &s->tmp, &s->ws,
                            &s->coef, &s->M2, &s->Qbar, &s->bad, &s->sketch.omega, &s->sketch.ws, &s->sketch.tail,
                            &s->gram.partial, &s->form.Z, &s->form.partial, &s->form.Wg,
                            &s->form.eigw, &s->form.eigv, &s->form.Wk, &s->form.S, &s->form.U,
                            &s->form.V})
      b += d->bytes;
    *bytes = b;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_set_seed(mcnla_solver s, uint64_t seed) {
  return guard([&] {
    MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "solver is null");
    s->cfg.seed = seed;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_solve(mcnla_solver s, const void* A, int64_t lda, int a_dtype,
                                int a_order, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "solver is null");
    check_a(s->prm, A, lda, a_dtype, a_order);
    const auto& p = s->prm;
    const auto& c = s->cfg;
    const int l = static_cast<int>(p.rank + p.over_rank);
    hipStream_t st = S(stream);
    SolveCtx cx;
    solve_begin(s, A, lda, a_dtype, a_order, st, cx, true);
    // 3. integrate on Bs, then Qbar = Y * (T * coef) in one pass over Y
    if (c.graph && !s->graphs) s->graphs = std::make_unique<IterationGraphs>();
    s->iteration = run_integration(c.integrator, s->Bs.as<double>(), p.num_sketch, l, c.maxiter,
                                   c.tol, s->coef.as<double>(), s->ws, st, &cx.conv, &cx.ok,
                                   s->graphs.get());
    s->converged = cx.conv;
    solve_finish(s, A, lda, a_dtype, a_order, st, cx, true);
    int bad = 0;
    read_bad(s, &bad, st);
    MCNLA_HIP(hipStreamSynchronize(st));
    return solve_status(s, cx, bad);
  });
}

mcnla_status mcnla_solver_solve_host(mcnla_solver s, const void* A, int64_t lda, int a_dtype,
                                     int a_order, int64_t chunk, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "solver is null");
    check_a(s->prm, A, lda, a_dtype, a_order);
    MCNLA_REQUIRE(chunk >= 0, MCNLA_ERR_INVALID, "chunk must be >= 0");
    const auto& p = s->prm;
    const auto& c = s->cfg;
    const int l = static_cast<int>(p.rank + p.over_rank);
    const int64_t nl = l * p.num_sketch;
    const bool cm = a_order == MCNLA_COL_MAJOR;
    const size_t esz = a_dtype == MCNLA_F32 ? sizeof(float) : sizeof(double);
    // A travels as `lines` lines of `len` elements: columns of a column-major A, rows of a
    // row-major one; its device copy is dense (leading dimension len)
    const int64_t len = cm ? p.nrow : p.ncol, lines = cm ? p.ncol : p.nrow;
    s->Acopy.ensure(static_cast<size_t>(len) * static_cast<size_t>(lines) * esz);
    if (!s->copy_stream) MCNLA_HIP(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    constexpr int64_t kMaxChunks = 64;
    const int64_t line_bytes = len * static_cast<int64_t>(esz);
    int64_t w = chunk > 0 ? chunk : std::max<int64_t>(1, (int64_t(1) << 30) / line_bytes);
    w = std::max(w, ceil_div(lines, kMaxChunks));
    const int64_t nch = ceil_div(lines, w);
    while (static_cast<int64_t>(s->chunk_ev.size()) < nch) {
      hipEvent_t e = nullptr;
      MCNLA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      s->chunk_ev.push_back(e);
    }
    hipStream_t st = S(stream);
    char* dA = s->Acopy.as<char>();
    SolveCtx cx;
    s->bind_res();
    s->timer.start(st);
    // The copy stream starts after what the caller enqueued on `st` before this call.
    MCNLA_HIP(hipEventRecord(s->chunk_ev[0], st));
    MCNLA_HIP(hipStreamWaitEvent(s->copy_stream, s->chunk_ev[0], 0));
    // On the sketch route in FP64, chunk k's share of Y = A Omega runs as soon as the chunk has
    // landed, beside the copy of the next one: column chunks add A(:, K_k) Omega(K_k, :) into Y
    // (the chunked la::mm with beta = 1 of col_block_gaussian_projection_sketcher_gpu.hpp:108-134),
    // row chunks write their rows of Y.  Otherwise the copy completes first and the solve runs
    // on the copy.
    const bool chunked = s->route == MCNLA_ROUTE_SKETCH && c.precision == MCNLA_PREC_F64;
    if (chunked)
      s->sketch.generate(p, c.seed, c.omega_mode, c.omega_threads, col_block_chain(c.sketcher),
                         st);
    for (int64_t k = 0; k < nch; ++k) {
      const int64_t l0 = k * w, l1 = std::min(lines, l0 + w);
      MCNLA_HIP(hipMemcpy2DAsync(dA + l0 * line_bytes, line_bytes,
                                 static_cast<const char*>(A) + l0 * lda * esz, lda * esz,
                                 line_bytes, l1 - l0, hipMemcpyHostToDevice, s->copy_stream));
      MCNLA_HIP(hipEventRecord(s->chunk_ev[k], s->copy_stream));
      if (!chunked) continue;
      MCNLA_HIP(hipStreamWaitEvent(st, s->chunk_ev[k], 0));
      GemmF64Args g;
      g.N = nl;
      g.a_f32 = (a_dtype == MCNLA_F32);
      g.ldb = nl;
      g.b_ncontig = true;
      g.ldc = nl;
      g.shape = GemmShape::kSquare;
      if (cm) {  // Y += A(:, l0:l1) Omega(l0:l1, :)
        g.M = p.nrow;
        g.K = l1 - l0;
        g.A = dA + l0 * line_bytes;
        g.lda = len;
        g.a_mcontig = true;
        g.B = s->sketch.omega.as<double>() + l0 * nl;
        g.C = s->Y.as<double>();
        g.beta = k == 0 ? 0.0 : 1.0;
      } else {  // Y(l0:l1, :) = A(l0:l1, :) Omega
        g.M = l1 - l0;
        g.K = p.ncol;
        g.A = dA + l0 * line_bytes;
        g.lda = len;
        g.a_mcontig = false;
        g.B = s->sketch.omega.p;
        g.C = s->Y.as<double>() + l0 * nl;
      }
      gemm_f64(g, st);
    }
    if (chunked) {
      cx.ldar = p.ncol;
      cx.rotate = !is_combining(c.integrator);
    } else {
      MCNLA_HIP(hipStreamWaitEvent(st, s->chunk_ev[nch - 1], 0));
      solve_sketch(s, dA, len, a_dtype, a_order, st, cx);
    }
    s->timer.mark(st);
    solve_orthogonalize(s, st, cx);
    s->timer.mark(st);
    if (c.graph && !s->graphs) s->graphs = std::make_unique<IterationGraphs>();
    s->iteration = run_integration(c.integrator, s->Bs.as<double>(), p.num_sketch, l, c.maxiter,
                                   c.tol, s->coef.as<double>(), s->ws, st, &cx.conv, &cx.ok,
                                   s->graphs.get());
    s->converged = cx.conv;
    solve_finish(s, dA, len, a_dtype, a_order, st, cx, true);
    int bad = 0;
    read_bad(s, &bad, st);
    MCNLA_HIP(hipStreamSynchronize(st));
    return solve_status(s, cx, bad);
  });
}

mcnla_status mcnla_solver_solve_batch(int count, const mcnla_solver* solvers,
                                      const void* const* A, const int64_t* lda,
                                      const int* a_dtype, const int* a_order,
                                      mcnla_status* status, void* stream) {
  return guard([&] {
    solve_batch(count, solvers, A, lda, a_dtype, a_order, status, nullptr, S(stream));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_results(mcnla_solver s, const double** Sv, const double** U,
                                  const double** V, const double** Qbar) {
  return guard([&] {
    MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "solver is null");
    MCNLA_REQUIRE(s->computed, MCNLA_ERR_STATE, "solver has not run");
    if (Sv) *Sv = s->form.S.as<double>();
    if (U) *U = s->form.U.as<double>();
    if (V)
      *V = (s->cfg.jobv && s->cfg.former != MCNLA_FORM_ROW_BLOCK_SYMMETRIC) ? s->form.V.as<double>()
                                                                              : nullptr;
    if (Qbar) *Qbar = s->Qbar.as<double>();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_iteration(mcnla_solver s, int64_t* iteration) {
  return guard([&] {
    MCNLA_REQUIRE(s && iteration, MCNLA_ERR_INVALID, "null argument");
    *iteration = s->iteration;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_stage_times(mcnla_solver s, double times[4], double* sketch_kernel_s) {
  return guard([&] {
    MCNLA_REQUIRE(s && times, MCNLA_ERR_INVALID, "null argument");
    MCNLA_REQUIRE(s->computed, MCNLA_ERR_STATE, "solver has not run");
    const auto v = s->timer.parts();
    for (int i = 0; i < 4; ++i) times[i] = i < static_cast<int>(v.size()) ? v[i] : 0.0;
    if (sketch_kernel_s) *sketch_kernel_s = s->sketch.gemm_seconds();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_kernel_time(mcnla_solver s, int kernel, double* seconds) {
  return guard([&] {
    MCNLA_REQUIRE(s && seconds, MCNLA_ERR_INVALID, "null argument");
    MCNLA_REQUIRE(kernel >= MCNLA_KERNEL_SKETCH && kernel <= MCNLA_KERNEL_FORMER_Z,
                  MCNLA_ERR_INVALID, "unknown kernel");
    MCNLA_REQUIRE(s->computed, MCNLA_ERR_STATE, "solver has not run");
    *seconds = kernel == MCNLA_KERNEL_SKETCH ? s->sketch.gemm_seconds()
               : kernel == MCNLA_KERNEL_GRAM ? s->gram.gemm_seconds()
                                             : s->form.z_seconds();
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_kernel_times(mcnla_solver s, double times[2]) {
  return guard([&] {
    MCNLA_REQUIRE(s && times, MCNLA_ERR_INVALID, "null argument");
    MCNLA_REQUIRE(s->computed, MCNLA_ERR_STATE, "solver has not run");
    times[0] = s->sketch.gemm_seconds();
    times[1] = s->gram.gemm_seconds();
    return MCNLA_OK;
  });
}

void mcnla_tuning_default(mcnla_tuning* t) {
  if (t) *t = kDefaultTuning;
}

mcnla_status mcnla_solver_get_tuning(mcnla_solver s, mcnla_tuning* t) {
  return guard([&] {
    MCNLA_REQUIRE(s && t, MCNLA_ERR_INVALID, "null argument");
    *t = s->tune;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_solver_autotune(mcnla_solver s, const void* A, int64_t lda, int a_dtype,
                                   int a_order, int reps, void* stream, mcnla_tuning* out) {
  return guard([&] {
    MCNLA_REQUIRE(s, MCNLA_ERR_INVALID, "solver is null");
    MCNLA_REQUIRE(reps >= 1 && reps <= 64, MCNLA_ERR_INVALID, "reps must be in [1, 64]");
    MCNLA_REQUIRE(!s->G.borrowed, MCNLA_ERR_STATE,
                  "a solver carved from a reserved region cannot be tuned (its split-K scratch "
                  "cannot grow)");
    check_a(s->prm, A, lda, a_dtype, a_order);
    // One knob per GEMM of the route, with the solve stages ([sketch, orthogonalize, integrate,
    // form]) that hold it and its bit in `kernels` (the alternative tile shape).
    struct Knob {
      int64_t mcnla_tuning::*field;
      std::vector<int64_t> candidates;
      unsigned stages;
      int64_t kernel_bit;
    };
    std::vector<Knob> knobs;
    if (s->route == MCNLA_ROUTE_GRAM_OF_A) {
      knobs = {{&mcnla_tuning::gram, {1152, 2304, 4608}, 1u, kKernGram},
               {&mcnla_tuning::h_omega, {1152, 2304, 4608}, 2u, kKernHOmega},
               {&mcnla_tuning::omega_p, {256, 512, 1024, 2048}, 2u, kKernOmegaP},
               {&mcnla_tuning::n_by_l, {128, 256, 512, 1024, 2048}, 12u, kKernNbyL}};
    } else {
      knobs = {{&mcnla_tuning::gram, {576, 1152, 2304, 4608, 9216}, 2u, kKernGram},
               {&mcnla_tuning::former, {1024, 2048, 4096, 8192}, 8u, kKernFormer},
               {&mcnla_tuning::qbar, {512, 1024, 2048, 3072, 4096}, 4u, kKernQbar}};
    }
    auto solve = [&] {
      const mcnla_status r = mcnla_solver_solve(s, A, lda, a_dtype, a_order, stream);
      MCNLA_REQUIRE(r == MCNLA_OK || r == MCNLA_ERR_NOT_CONVERGED, r, get_error());
    };
    // best time of `reps` solves over the knob's stages, after one untimed solve
    auto measure = [&](unsigned stages) {
      solve();
      double best = 0.0;
      for (int i = 0; i < reps; ++i) {
        solve();
        const std::vector<double> v = s->timer.parts();
        double t = 0.0;
        for (int j = 0; j < 4 && j < static_cast<int>(v.size()); ++j)
          if (stages & (1u << j)) t += v[j];
        if (i == 0 || t < best) best = t;
      }
      return best;
    };
    mcnla_tuning t = s->tune;
    for (const Knob& k : knobs) {
      int64_t best_v = t.*k.field;
      double best_t = measure(k.stages);
      for (int64_t v : k.candidates) {
        if (v == best_v) continue;
        mcnla_tuning c = t;
        c.*k.field = v;
        apply_tuning(s, c);
        const double ct = measure(k.stages);
        if (ct < 0.99 * best_t) {
          best_t = ct;
          best_v = v;
        }
      }
      t.*k.field = best_v;
      apply_tuning(s, t);
      // then the GEMM's other kernel (tile shape) at the chosen target
      mcnla_tuning c = t;
      c.kernels ^= k.kernel_bit;
      apply_tuning(s, c);
      if (measure(k.stages) < 0.99 * best_t) t = c;
      apply_tuning(s, t);
    }
    solve();  // leave the results of the tuned plan
    store_tuning(tune_key(s->prm, s->route, s->cfg.precision), t);
    if (out) *out = t;
    return MCNLA_OK;
  });
}

mcnla_status mcnla_tune_set(const mcnla_params* params, const mcnla_solver_config* cfg,
                            const mcnla_tuning* t) {
  return guard([&] {
    MCNLA_REQUIRE(params && cfg && t, MCNLA_ERR_INVALID, "null argument");
    check_params(params);
    MCNLA_REQUIRE(valid_tuning(*t), MCNLA_ERR_INVALID, "tuning targets must be in [1, 2^20]");
    store_tuning(tune_key(*params, resolve_route(*params, *cfg), cfg->precision), *t);
    return MCNLA_OK;
  });
}

mcnla_status mcnla_tune_get(const mcnla_params* params, const mcnla_solver_config* cfg,
                            mcnla_tuning* t, int* found) {
  return guard([&] {
    MCNLA_REQUIRE(params && cfg && t, MCNLA_ERR_INVALID, "null argument");
    check_params(params);
    bool f = false;
    *t = lookup_tuning(tune_key(*params, resolve_route(*params, *cfg), cfg->precision), &f);
    if (found) *found = f ? 1 : 0;
    return MCNLA_OK;
  });
}

void mcnla_tune_clear(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_table().clear();
}

mcnla_status mcnla_tune_save(const char* path) {
  return guard([&] {
    MCNLA_REQUIRE(path, MCNLA_ERR_INVALID, "null argument");
    std::string text;
    {
      std::lock_guard<std::mutex> lock(g_tune_mu);
      for (const auto& e : tune_table()) {
        const TuneKey& k = e.first;
        const mcnla_tuning& t = e.second;
        char line[256];
        std::snprintf(line, sizeof line,
                      "%lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld\n",
                      (long long)k.v[0], (long long)k.v[1], (long long)k.v[2], (long long)k.v[3],
                      (long long)k.v[4], (long long)k.v[5], (long long)t.gram,
                      (long long)t.former, (long long)t.h_omega, (long long)t.omega_p,
                      (long long)t.n_by_l, (long long)t.qbar, (long long)t.kernels);
        text += line;
      }
    }
    FILE* f = std::fopen(path, "w");
    MCNLA_REQUIRE(f, MCNLA_ERR_INVALID, std::string("cannot open ") + path);
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    MCNLA_REQUIRE(std::fclose(f) == 0 && ok, MCNLA_ERR_INVALID,
                  std::string("cannot write ") + path);
    return MCNLA_OK;
  });
}

mcnla_status mcnla_tune_load(const char* path) {
  return guard([&] {
    MCNLA_REQUIRE(path, MCNLA_ERR_INVALID, "null argument");
    FILE* f = std::fopen(path, "r");
    MCNLA_REQUIRE(f, MCNLA_ERR_INVALID, std::string("cannot open ") + path);
    std::vector<std::pair<TuneKey, mcnla_tuning>> entries;
    char line[512];
    int lineno = 0;
    std::string bad;
    while (bad.empty() && std::fgets(line, sizeof line, f)) {
      ++lineno;
      long long v[13] = {0};
      char extra;
      const int got = std::sscanf(line,
                                  "%lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %c",
                                  &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8],
                                  &v[9], &v[10], &v[11], &v[12], &extra);
      bool blank = true;
      for (const char* c = line; *c; ++c)
        if (*c != ' ' && *c != '\t' && *c != '\n' && *c != '\r') blank = false;
      if (blank) continue;
      TuneKey k{{v[0], v[1], v[2], v[3], v[4], v[5]}};
      const mcnla_tuning t{v[6], v[7], v[8], v[9], v[10], v[11], v[12]};
      const bool key_ok = (got == 12 || got == 13) &&
                          (k.v[0] == MCNLA_ROUTE_SKETCH || k.v[0] == MCNLA_ROUTE_GRAM_OF_A) &&
                          k.v[2] >= 1 && k.v[3] >= 1 && k.v[4] >= 1 && k.v[5] >= 1;
      if (!key_ok || !valid_tuning(t)) bad = path + std::string(":") + std::to_string(lineno);
      else entries.emplace_back(k, t);
    }
    std::fclose(f);
    MCNLA_REQUIRE(bad.empty(), MCNLA_ERR_INVALID, "malformed tuning entry at " + bad);
    for (const auto& e : entries) store_tuning(e.first, e.second);
    return MCNLA_OK;
  });
}

mcnla_status mcnla_synth_lowrank(int64_t m, int64_t n, int64_t r, const double* U, const double* V,
                                 const double* sigma, double noise, uint64_t seed, void* A,
                                 int64_t lda, int a_dtype, int a_order, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(m >= 1 && n >= 1 && r >= 0 && A, MCNLA_ERR_INVALID, "bad synth arguments");
    MCNLA_REQUIRE(r == 0 || (U && V && sigma), MCNLA_ERR_INVALID, "null factors");
    MCNLA_REQUIRE(lda >= (a_order == MCNLA_COL_MAJOR ? m : n), MCNLA_ERR_INVALID, "lda too small");
    hipStream_t st = S(stream);
    if (r <= 64) {
      synth_lowrank(m, n, r, U, V, sigma, noise, seed, A, lda, a_dtype == MCNLA_F32,
                    a_order == MCNLA_COL_MAJOR, st);
      return MCNLA_OK;
    }
    // High rank: T = U (V diag(sigma))' on the FP64 MFMA GEMM, then noise and storage order.
    DevBuf Vs, Tm;
    Vs.ensure(static_cast<size_t>(n) * r * sizeof(double));
    Tm.ensure(static_cast<size_t>(m) * n * sizeof(double));
    MCNLA_HIP(hipMemcpyAsync(Vs.p, V, static_cast<size_t>(n) * r * sizeof(double),
                             hipMemcpyDeviceToDevice, st));
    scale_columns(Vs.as<double>(), n, static_cast<int>(r), sigma, false, st);
    GemmF64Args g;
    g.M = m;
    g.N = n;
    g.K = r;
    g.A = U;
    g.lda = r;
    g.a_mcontig = false;
    g.B = Vs.p;
    g.ldb = r;
    g.b_ncontig = false;
    g.C = Tm.as<double>();
    g.ldc = n;
    g.shape = GemmShape::kSquare;
    gemm_f64(g, st);
    synth_finish(m, n, Tm.as<double>(), noise, seed, A, lda, a_dtype == MCNLA_F32,
                 a_order == MCNLA_COL_MAJOR, st);
    MCNLA_HIP(hipStreamSynchronize(st));  // the temporaries are freed on return
    return MCNLA_OK;
  });
}

mcnla_status mcnla_synth_reflected(int64_t m, int64_t n, const double* sigma, const double* u,
                                   const double* v, void* A, int64_t lda, int a_dtype,
                                   int a_order, void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(m >= 1 && n >= 1 && n <= m && A && sigma && u && v, MCNLA_ERR_INVALID,
                  "bad synth arguments (need 1 <= n <= m and non-null vectors)");
    MCNLA_REQUIRE(a_dtype == MCNLA_F64 || a_dtype == MCNLA_F32, MCNLA_ERR_INVALID, "A dtype");
    MCNLA_REQUIRE(a_order == MCNLA_COL_MAJOR || a_order == MCNLA_ROW_MAJOR, MCNLA_ERR_INVALID,
                  "A order");
    MCNLA_REQUIRE(lda >= (a_order == MCNLA_COL_MAJOR ? m : n), MCNLA_ERR_INVALID, "lda too small");
    synth_reflected(m, n, sigma, u, v, A, lda, a_dtype == MCNLA_F32, a_order == MCNLA_COL_MAJOR,
                    S(stream));
    return MCNLA_OK;
  });
}

mcnla_status mcnla_fill_gaussian(double* out, int64_t count, uint64_t seed, double scale,
                                 void* stream) {
  return guard([&] {
    MCNLA_REQUIRE(count >= 0 && (count == 0 || out), MCNLA_ERR_INVALID, "bad fill arguments");
    gaussian_fill(out, count, seed, scale, S(stream));
    return MCNLA_OK;
  });
}

}  // extern "C"
