// Header-only C++ host layer of mcnla_amd: the reference's mcnla::isvd stage-plugin API
// (emfomy/mcnla 0.7.2, include/mcnla/isvd/**) over the C ABI of libmcnla_amd.so.
//
// Drop-in surface, same names and call order as the reference drivers
// (src/isvd_dense_cb.cpp:118-243):
//   mcnla::isvd::Parameters<double> parameters(root, comm);
//   parameters.setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N); parameters.sync();
//   mcnla::isvd::ColBlockGaussianProjectionSketcher<double> sketcher(parameters);
//   mcnla::isvd::RowBlockGramianOrthogonalizer<double>      orthogonalizer(parameters);
//   mcnla::isvd::RowBlockKolmogorovNagumoIntegrator<double> integrator(parameters);
//   mcnla::isvd::SvdFormer<double, true>                    former(parameters);
//   sketcher.setSeed(seed); ...initialize();
//   sketcher(A, Q); orthogonalizer(Q); integrator(Q, Qbar); former(A, Qbar);
//   former.vectorS(); former.matrixU(); former.matrixV(); integrator.iteration(); x.time();
//
// Containers keep a host copy (element access like the reference's DenseMatrix) and a device
// mirror in HBM; a stage call moves data only when the side it reads is stale.  All compute runs
// on the GPU; there is no CPU path.  Errors: the reference is noexcept and asserts; here a failed
// C-ABI call throws mcnla::Error (the drivers catch it and exit non-zero).
#ifndef MCNLA_AMD_ISVD_HPP_
#define MCNLA_AMD_ISVD_HPP_

#include <cstdint>
#include <cstring>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "isvd_capi.h"

namespace mcnla {

using index_t = int64_t;

struct Error : std::runtime_error {
  mcnla_status status;
  Error(mcnla_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(mcnla_status s, const char* where) {
  if (s != MCNLA_OK) throw Error(s, std::string(where) + ": " + mcnla_last_error());
}

// The reference's MPI plumbing reduces to one rank on one GPU.
using MPI_Comm_t = int;
constexpr MPI_Comm_t kCommWorld = 0;

enum class Trans { NORMAL, TRANS };

namespace matrix {

// Host buffer + lazily synchronized device mirror.  Copies share storage (the reference's
// shallow-copy semantics, include/mcnla/core/matrix/kit/array.hpp:47).
template <typename T>
class Storage {
 public:
  explicit Storage(size_t n = 0) : s_(std::make_shared<State>(n)) {}
  size_t size() const { return s_->host.size(); }
  T* host() {
    pull();
    s_->where = kHost;
    return s_->host.data();
  }
  const T* host() const {
    pull();
    return s_->host.data();
  }
  const T* device() const {
    push();
    return static_cast<const T*>(s_->dev);
  }
  T* deviceMutable() {
    push();
    s_->where = kDevice;
    return static_cast<T*>(s_->dev);
  }

 private:
  enum Where { kBoth, kHost, kDevice };
  struct State {
    std::vector<T> host;
    void* dev = nullptr;
    Where where = kHost;
    explicit State(size_t n) : host(n) {}
    ~State() {
      if (dev) mcnla_free(dev);
    }
  };
  void ensure() const {
    if (!s_->dev && !s_->host.empty())
      check(mcnla_malloc(&s_->dev, s_->host.size() * sizeof(T)), "mcnla_malloc");
  }
  void push() const {
    ensure();
    if (s_->where == kHost) {
      if (!s_->host.empty())
        check(mcnla_memcpy_htod(s_->dev, s_->host.data(), s_->host.size() * sizeof(T)), "htod");
      s_->where = kBoth;
    }
  }
  void pull() const {
    if (s_->where == kDevice) {
      check(mcnla_memcpy_dtoh(s_->host.data(), s_->dev, s_->host.size() * sizeof(T)), "dtoh");
      s_->where = kBoth;
    }
  }
  std::shared_ptr<State> s_;
};

template <typename T>
class DenseVector {
 public:
  DenseVector() = default;
  explicit DenseVector(index_t len) : len_(len), st_(static_cast<size_t>(len)) {}
  void reconstruct(index_t len) { *this = DenseVector(len); }
  index_t len() const { return len_; }
  T& operator()(index_t i) { return st_.host()[i]; }
  T operator()(index_t i) const { return st_.host()[i]; }
  T* valPtr() { return st_.host(); }
  const T* valPtr() const { return st_.host(); }
  Storage<T>& storage() { return st_; }
  const Storage<T>& storage() const { return st_; }

 private:
  index_t len_ = 0;
  Storage<T> st_;
};

// Dense m x n matrix; ROW = row-major (the reference's DenseMatrixRowMajor).  The leading
// dimension equals the row length (row-major) or the column length (column-major).
template <typename T, bool ROW>
class DenseMatrix {
 public:
  DenseMatrix() = default;
  DenseMatrix(index_t m, index_t n) : m_(m), n_(n), st_(static_cast<size_t>(m * n)) {}
  void reconstruct(index_t m, index_t n) { *this = DenseMatrix(m, n); }
  bool isEmpty() const { return m_ * n_ == 0; }
  index_t nrow() const { return m_; }
  index_t ncol() const { return n_; }
  index_t pitch() const { return ROW ? n_ : m_; }
  std::tuple<index_t, index_t> sizes() const { return std::make_tuple(m_, n_); }
  T& operator()(index_t i, index_t j) { return st_.host()[ROW ? i * n_ + j : i + j * m_]; }
  T operator()(index_t i, index_t j) const { return st_.host()[ROW ? i * n_ + j : i + j * m_]; }
  T* valPtr() { return st_.host(); }
  const T* valPtr() const { return st_.host(); }
  Storage<T>& storage() { return st_; }
  const Storage<T>& storage() const { return st_; }
  static constexpr int order() { return ROW ? MCNLA_ROW_MAJOR : MCNLA_COL_MAJOR; }
  static constexpr int dtype() { return sizeof(T) == 8 ? MCNLA_F64 : MCNLA_F32; }

 private:
  index_t m_ = 0, n_ = 0;
  Storage<T> st_;
};

template <typename T> using DenseMatrixRowMajor = DenseMatrix<T, true>;
template <typename T> using DenseMatrixColMajor = DenseMatrix<T, false>;

// The col-block collection: ONE row-major m x (n*N) buffer, matrix i in columns [i*n, (i+1)*n)
// (include/mcnla/core/matrix/collection/dense/dense_matrix_collection_col_block.hpp:329-345).
template <typename T>
class DenseMatrixCollectionColBlockRowMajor {
 public:
  DenseMatrixCollectionColBlockRowMajor() = default;
  DenseMatrixCollectionColBlockRowMajor(index_t m, index_t n, index_t N)
      : m_(m), n_(n), N_(N), unfold_(m, n * N) {}
  void reconstruct(index_t m, index_t n, index_t N) { *this = DenseMatrixCollectionColBlockRowMajor(m, n, N); }
  bool isEmpty() const { return m_ * n_ * N_ == 0; }
  index_t nrow() const { return m_; }
  index_t ncol() const { return n_; }
  index_t nmat() const { return N_; }
  std::tuple<index_t, index_t, index_t> sizes() const { return std::make_tuple(m_, n_, N_); }
  // element (i, j) of matrix k
  T& operator()(index_t i, index_t j, index_t k) { return unfold_(i, k * n_ + j); }
  T operator()(index_t i, index_t j, index_t k) const { return unfold_(i, k * n_ + j); }
  DenseMatrixRowMajor<T>& unfold() { return unfold_; }
  const DenseMatrixRowMajor<T>& unfold() const { return unfold_; }

 private:
  index_t m_ = 0, n_ = 0, N_ = 0;
  DenseMatrixRowMajor<T> unfold_;
};

}  // namespace matrix

using matrix::DenseMatrixCollectionColBlockRowMajor;
using matrix::DenseMatrixColMajor;
using matrix::DenseMatrixRowMajor;
using matrix::DenseVector;

namespace isvd {

// isvd::Parameters (include/mcnla/isvd/core/parameters.hh:31-119).  One rank: the per-rank
// sizes equal the global ones.
template <typename _Val>
class Parameters {
 public:
  const MPI_Comm_t mpi_comm;
  const int mpi_size = 1;
  const int mpi_root;
  const int mpi_rank = 0;

  Parameters(int root = 0, MPI_Comm_t comm = kCommWorld) : mpi_comm(comm), mpi_root(root) {}
  void sync() { synchronized_ = true; }
  bool isSynchronized() const { return synchronized_; }
  index_t nrow() const { return p_.nrow; }
  index_t ncol() const { return p_.ncol; }
  index_t nrowRank() const { return p_.nrow; }
  index_t ncolRank() const { return p_.ncol; }
  index_t nrowEach() const { return p_.nrow; }
  index_t ncolEach() const { return p_.ncol; }
  index_t nrowTotal() const { return p_.nrow; }
  index_t ncolTotal() const { return p_.ncol; }
  index_t rank() const { return p_.rank; }
  index_t overRank() const { return p_.over_rank; }
  index_t dimSketch() const { return p_.rank + p_.over_rank; }
  index_t dimSketchEach() const { return dimSketch(); }
  index_t dimSketchTotal() const { return dimSketch() * p_.num_sketch; }
  index_t numSketch() const { return p_.num_sketch; }
  index_t numSketchEach() const { return p_.num_sketch; }

  template <class _Matrix>
  Parameters& setSize(const _Matrix& a) { return setSize(a.nrow(), a.ncol()); }
  Parameters& setSize(index_t nrow, index_t ncol) {
    p_.nrow = nrow;
    p_.ncol = ncol;
    synchronized_ = false;
    return *this;
  }
  Parameters& setRank(index_t k) { p_.rank = k; return *this; }
  Parameters& setOverRank(index_t p) { p_.over_rank = p; return *this; }
  Parameters& setNumSketch(index_t N) { p_.num_sketch = N; return *this; }
  Parameters& setNumSketchEach(index_t N) { p_.num_sketch = N; return *this; }

  DenseMatrixCollectionColBlockRowMajor<_Val> createCollectionQ() const {
    return DenseMatrixCollectionColBlockRowMajor<_Val>(nrow(), dimSketch(), numSketch());
  }
  DenseMatrixCollectionColBlockRowMajor<_Val> createCollectionQj() const { return createCollectionQ(); }
  DenseMatrixCollectionColBlockRowMajor<_Val> createCollectionQjp() const { return createCollectionQ(); }
  DenseMatrixRowMajor<_Val> createMatrixQbar() const { return DenseMatrixRowMajor<_Val>(nrow(), dimSketch()); }
  DenseMatrixRowMajor<_Val> createMatrixQbarj() const { return createMatrixQbar(); }
  DenseMatrixRowMajor<_Val> createMatrixU() const { return DenseMatrixRowMajor<_Val>(nrow(), rank()); }
  DenseMatrixRowMajor<_Val> createMatrixUj() const { return createMatrixU(); }
  DenseMatrixRowMajor<_Val> createMatrixV() const { return DenseMatrixRowMajor<_Val>(ncol(), rank()); }
  DenseMatrixRowMajor<_Val> createMatrixVj() const { return createMatrixV(); }

  const mcnla_params& c() const { return p_; }

 private:
  bool synchronized_ = false;
  mcnla_params p_{0, 0, 0, 12, 16};
};

// StageWrapper lifecycle (include/mcnla/isvd/core/stage_wrapper.hh:35-101) over a C-ABI handle.
class StageBase {
 public:
  StageBase(const StageBase&) = delete;
  StageBase& operator=(const StageBase&) = delete;
  virtual ~StageBase() {
    if (h_) mcnla_stage_destroy(h_);
  }
  bool isInitialized() const { return h_ != nullptr; }
  bool isComputed() const { return computed_; }
  double time() const {
    double t = 0.0;
    check(mcnla_stage_time(h_, &t, nullptr, 0, nullptr), "time");
    return t;
  }
  std::vector<double> times() const {
    double parts[8];
    int n = 0;
    check(mcnla_stage_time(h_, nullptr, parts, 8, &n), "times");
    return std::vector<double>(parts, parts + n);
  }
  double commTime() const { return 0.0; }  // no communication on one GPU
  const char* names() const { return h_ ? mcnla_stage_part_names(h_) : ""; }
  friend std::ostream& operator<<(std::ostream& os, const StageBase& s) {
    return os << (s.h_ ? mcnla_stage_name(s.h_) : "(uninitialized stage)");
  }

 protected:
  StageBase() = default;
  void reset(mcnla_stage h) {
    if (h_) mcnla_stage_destroy(h_);
    h_ = h;
    computed_ = false;
  }
  mcnla_stage h_ = nullptr;
  bool computed_ = false;
};

// ---------------------------------------------------------------------------------------------
// Sketchers (include/mcnla/isvd/sketcher/*.hh).
template <int KIND, typename _Val>
class GaussianSketcherImpl : public StageBase {
 public:
  explicit GaussianSketcherImpl(const Parameters<_Val>& p, index_t seed = 0, index_t exponent = 0)
      : params_(p), seed_(seed), exponent_(exponent) {}
  void initialize() {
    mcnla_stage h = nullptr;
    check(mcnla_sketcher_create(&params_.c(), KIND, precision_, &h), "sketcher");
    reset(h);
    check(mcnla_sketcher_set_seed(h_, static_cast<uint64_t>(seed_), omega_mode_, threads_), "setSeed");
    check(mcnla_sketcher_set_exponent(h_, exponent_), "setExponent");
  }
  // Power iterations (gaussian_projection_sketcher.hh:87; GaussianProjectionSketcher only).
  GaussianSketcherImpl& setExponent(index_t exponent) {
    exponent_ = exponent;
    if (h_) check(mcnla_sketcher_set_exponent(h_, exponent_), "setExponent");
    return *this;
  }
  index_t exponent() const { return exponent_; }
  GaussianSketcherImpl& setSeed(index_t seed) {
    seed_ = seed;
    if (h_) check(mcnla_sketcher_set_seed(h_, static_cast<uint64_t>(seed_), omega_mode_, threads_), "setSeed");
    return *this;
  }
  // Selects the Omega stream (MCNLA_OMEGA_PHILOX / MCNLA_OMEGA_REFERENCE) and, for the
  // reference stream, the OpenMP thread count of the run to reproduce.
  GaussianSketcherImpl& setOmega(int mode, int threads = 1) {
    omega_mode_ = mode;
    threads_ = threads;
    return setSeed(seed_);
  }
  GaussianSketcherImpl& setPrecision(int precision) {
    precision_ = precision;
    return *this;
  }
  index_t seed() const { return seed_; }
  template <class _Matrix>
  void operator()(const _Matrix& a, DenseMatrixCollectionColBlockRowMajor<_Val>& q) {
    if (!h_) initialize();
    check(mcnla_sketcher_run(h_, a.storage().device(), a.pitch(), _Matrix::dtype(), _Matrix::order(),
                             q.unfold().storage().deviceMutable(), nullptr),
          "sketcher");
    computed_ = true;
  }

 private:
  const Parameters<_Val>& params_;
  index_t seed_;
  index_t exponent_;
  int omega_mode_ = MCNLA_OMEGA_PHILOX, threads_ = 1, precision_ = MCNLA_PREC_F64;
};

template <typename _Val> using ColBlockGaussianProjectionSketcher =
    GaussianSketcherImpl<MCNLA_SKETCHER_COL_BLOCK_GAUSSIAN, _Val>;
template <typename _Val> using RowBlockGaussianProjectionSketcher =
    GaussianSketcherImpl<MCNLA_SKETCHER_ROW_BLOCK_GAUSSIAN, _Val>;
template <typename _Val> using GaussianProjectionSketcher =
    GaussianSketcherImpl<MCNLA_SKETCHER_GAUSSIAN, _Val>;

// Column sampling (column_sampling_sketcher.hh, row_block_column_sampling_sketcher.hh): Y's
// columns are columns of A drawn by the reference's seeded index stream; `threads` is the OpenMP
// thread count of the run whose stream should be reproduced.
template <int KIND, typename _Val>
class ColumnSamplingSketcherImpl : public StageBase {
 public:
  explicit ColumnSamplingSketcherImpl(const Parameters<_Val>& p, index_t seed = 0)
      : params_(p), seed_(seed) {}
  void initialize() {
    mcnla_stage h = nullptr;
    check(mcnla_sketcher_create(&params_.c(), KIND, MCNLA_PREC_F64, &h), "sketcher");
    reset(h);
    check(mcnla_sketcher_set_seed(h_, static_cast<uint64_t>(seed_), MCNLA_OMEGA_REFERENCE, threads_),
          "setSeed");
  }
  ColumnSamplingSketcherImpl& setSeed(index_t seed, int threads = 1) {
    seed_ = seed;
    threads_ = threads;
    if (h_)
      check(mcnla_sketcher_set_seed(h_, static_cast<uint64_t>(seed_), MCNLA_OMEGA_REFERENCE, threads_),
            "setSeed");
    return *this;
  }
  index_t seed() const { return seed_; }
  template <class _Matrix>
  void operator()(const _Matrix& a, DenseMatrixCollectionColBlockRowMajor<_Val>& q) {
    if (!h_) initialize();
    check(mcnla_sketcher_run(h_, a.storage().device(), a.pitch(), _Matrix::dtype(), _Matrix::order(),
                             q.unfold().storage().deviceMutable(), nullptr),
          "sketcher");
    computed_ = true;
  }

 private:
  const Parameters<_Val>& params_;
  index_t seed_;
  int threads_ = 1;
};

template <typename _Val> using ColumnSamplingSketcher =
    ColumnSamplingSketcherImpl<MCNLA_SKETCHER_COLUMN_SAMPLING, _Val>;
template <typename _Val> using RowBlockColumnSamplingSketcher =
    ColumnSamplingSketcherImpl<MCNLA_SKETCHER_ROW_BLOCK_COLUMN_SAMPLING, _Val>;

// ---------------------------------------------------------------------------------------------
// Orthogonalizers (include/mcnla/isvd/orthogonalizer/*gramian*.hh).
template <int KIND, typename _Val>
class GramianOrthogonalizerImpl : public StageBase {
 public:
  explicit GramianOrthogonalizerImpl(const Parameters<_Val>& p) : params_(p) {}
  void initialize() {
    mcnla_stage h = nullptr;
    check(mcnla_orthogonalizer_create(&params_.c(), KIND, MCNLA_PREC_F64, &h), "orthogonalizer");
    reset(h);
  }
  void operator()(DenseMatrixCollectionColBlockRowMajor<_Val>& q) {
    if (!h_) initialize();
    check(mcnla_orthogonalizer_run(h_, q.unfold().storage().deviceMutable(), nullptr), "orthogonalizer");
    computed_ = true;
  }

 private:
  const Parameters<_Val>& params_;
};

template <typename _Val> using GramianOrthogonalizer = GramianOrthogonalizerImpl<MCNLA_ORTH_GRAMIAN, _Val>;
template <typename _Val> using RowBlockGramianOrthogonalizer =
    GramianOrthogonalizerImpl<MCNLA_ORTH_ROW_BLOCK_GRAMIAN, _Val>;

// ---------------------------------------------------------------------------------------------
// Integrators (include/mcnla/isvd/integrator/*.hh); defaults maxiter 256, tol 1e-4 (KN) or
// 1e-3 (WY) as in row_block_kolmogorov_nagumo_integrator.hh:115-116 and
// row_block_wen_yin_integrator.hh:135-136.
template <int KIND, typename _Val>
class IntegratorImpl : public StageBase {
 public:
  static constexpr bool kWenYin = KIND == MCNLA_INT_ROW_BLOCK_WY || KIND == MCNLA_INT_ROW_BLOCK_GRAMIAN_WY;
  explicit IntegratorImpl(const Parameters<_Val>& p, index_t max_iteration = 256,
                          double tolerance = kWenYin ? 1e-3 : 1e-4)
      : params_(p), maxiter_(max_iteration), tol_(tolerance) {}
  void initialize() {
    mcnla_stage h = nullptr;
    check(mcnla_integrator_create(&params_.c(), KIND, MCNLA_PREC_F64, maxiter_, tol_, &h), "integrator");
    reset(h);
  }
  IntegratorImpl& setMaxIteration(index_t m) {
    maxiter_ = m;
    if (h_) initialize();
    return *this;
  }
  IntegratorImpl& setTolerance(double t) {
    tol_ = t;
    if (h_) initialize();
    return *this;
  }
  index_t maxIteration() const { return maxiter_; }
  double tolerance() const { return tol_; }
  index_t iteration() const { return iteration_; }
  bool converged() const { return converged_; }
  void operator()(const DenseMatrixCollectionColBlockRowMajor<_Val>& q, DenseMatrixRowMajor<_Val>& qbar) {
    if (!h_) initialize();
    const mcnla_status s = mcnla_integrator_run(h_, q.unfold().storage().device(),
                                                qbar.storage().deviceMutable(), nullptr);
    if (s != MCNLA_ERR_NOT_CONVERGED) check(s, "integrator");
    converged_ = (s == MCNLA_OK);
    int64_t it = 0;
    check(mcnla_integrator_iteration(h_, &it), "iteration");
    iteration_ = it;
    computed_ = true;
  }

 private:
  const Parameters<_Val>& params_;
  index_t maxiter_;
  double tol_;
  index_t iteration_ = 0;
  bool converged_ = false;
};

template <typename _Val> using KolmogorovNagumoIntegrator = IntegratorImpl<MCNLA_INT_KN, _Val>;
template <typename _Val> using RowBlockKolmogorovNagumoIntegrator = IntegratorImpl<MCNLA_INT_ROW_BLOCK_KN, _Val>;
template <typename _Val> using RowBlockGramianKolmogorovNagumoIntegrator =
    IntegratorImpl<MCNLA_INT_ROW_BLOCK_GRAMIAN_KN, _Val>;
template <typename _Val> using RowBlockWenYinIntegrator = IntegratorImpl<MCNLA_INT_ROW_BLOCK_WY, _Val>;
template <typename _Val> using RowBlockGramianWenYinIntegrator =
    IntegratorImpl<MCNLA_INT_ROW_BLOCK_GRAMIAN_WY, _Val>;
// The combining integrators do not iterate (maxIteration/tolerance are unused, iteration() is 0);
// the extrinsic mean's (Qj, Q, Qbar) call form is accepted with Qj = Q on one rank
// (row_block_extrinsic_mean_integrator.hh:109-111).
template <typename _Val> using RowBlockReductionIntegrator =
    IntegratorImpl<MCNLA_INT_ROW_BLOCK_REDUCTION, _Val>;
template <typename _Val>
class RowBlockExtrinsicMeanIntegrator
    : public IntegratorImpl<MCNLA_INT_ROW_BLOCK_EXTRINSIC_MEAN, _Val> {
  using Base = IntegratorImpl<MCNLA_INT_ROW_BLOCK_EXTRINSIC_MEAN, _Val>;

 public:
  using Base::Base;
  using Base::operator();
  void operator()(const DenseMatrixCollectionColBlockRowMajor<_Val>& qj,
                  const DenseMatrixCollectionColBlockRowMajor<_Val>& /*q*/,
                  DenseMatrixRowMajor<_Val>& qbar) {
    Base::operator()(qj, qbar);
  }
};

// ---------------------------------------------------------------------------------------------
// Formers (include/mcnla/isvd/former/*.hh).  The SVD former returns U column-major
// (svd_former.hh:108); the Gramian formers row-major.  V is row-major n x k.
template <int KIND, typename _Val, bool _jobv>
class FormerImpl : public StageBase {
 public:
  using UType = typename std::conditional<KIND == MCNLA_FORM_SVD, DenseMatrixColMajor<_Val>,
                                          DenseMatrixRowMajor<_Val>>::type;
  explicit FormerImpl(const Parameters<_Val>& p) : params_(p) {}
  void initialize() {
    mcnla_stage h = nullptr;
    check(mcnla_former_create(&params_.c(), KIND, _jobv ? 1 : 0, MCNLA_PREC_F64, &h), "former");
    reset(h);
  }
  template <class _Matrix>
  void operator()(const _Matrix& a, const DenseMatrixRowMajor<_Val>& qbar) {
    if (!h_) initialize();
    check(mcnla_former_run(h_, a.storage().device(), a.pitch(), _Matrix::dtype(), _Matrix::order(),
                           qbar.storage().device(), nullptr),
          "former");
    const double *S = nullptr, *U = nullptr, *V = nullptr;
    check(mcnla_former_results(h_, &S, &U, &V), "results");
    const index_t m = params_.nrow(), n = params_.ncol(), k = params_.rank();
    s_.reconstruct(k);
    u_.reconstruct(m, k);
    check(mcnla_memcpy_dtoh(s_.valPtr(), S, k * sizeof(double)), "S");
    check(mcnla_memcpy_dtoh(u_.valPtr(), U, m * k * sizeof(double)), "U");
    if (_jobv && V != nullptr) {
      v_.reconstruct(n, k);
      check(mcnla_memcpy_dtoh(v_.valPtr(), V, n * k * sizeof(double)), "V");
    }
    computed_ = true;
  }
  const DenseVector<_Val>& vectorS() const { return s_; }
  const UType& matrixU() const { return u_; }
  const UType& matrixUj() const { return u_; }
  const DenseMatrixRowMajor<_Val>& matrixV() const { return v_; }
  const DenseMatrixRowMajor<_Val>& matrixVj() const { return v_; }

 private:
  const Parameters<_Val>& params_;
  DenseVector<_Val> s_;
  UType u_;
  DenseMatrixRowMajor<_Val> v_;
};

template <typename _Val, bool _jobv = true> using SvdFormer = FormerImpl<MCNLA_FORM_SVD, _Val, _jobv>;
template <typename _Val, bool _jobv = true> using GramianFormer = FormerImpl<MCNLA_FORM_GRAMIAN, _Val, _jobv>;
template <typename _Val, bool _jobv = true> using ColBlockGramianFormer =
    FormerImpl<MCNLA_FORM_COL_BLOCK_GRAMIAN, _Val, _jobv>;
template <typename _Val, bool _jobv = true> using RowBlockGramianFormer =
    FormerImpl<MCNLA_FORM_ROW_BLOCK_GRAMIAN, _Val, _jobv>;
// Square symmetric A: vectorS() = the k largest Ritz values (ascending), matrixUj() = Qbar W;
// V is not formed (row_block_symmetric_former.hpp:98-131).
template <typename _Val, bool _jobv = true> using RowBlockSymmetricFormer =
    FormerImpl<MCNLA_FORM_ROW_BLOCK_SYMMETRIC, _Val, _jobv>;

// ---------------------------------------------------------------------------------------------
// Converters (include/mcnla/isvd/converter/*.hh): with every row and every sketch on one GPU
// they are the identity (a device copy only when given two different containers).
template <typename _Val>
class ConverterImpl {
 public:
  explicit ConverterImpl(const Parameters<_Val>&) {}
  void initialize() {}
  double time() const { return 0.0; }
  template <class _In, class _Out>
  void operator()(const _In& in, _Out& out) {
    copy(in, out);
  }

 private:
  static void copy(const DenseMatrixCollectionColBlockRowMajor<_Val>& in,
                   DenseMatrixCollectionColBlockRowMajor<_Val>& out) {
    if (&in != &out) out = in;
  }
  template <class _M1, class _M2>
  static void copy(const _M1& in, _M2& out) {
    if (static_cast<const void*>(&in) != static_cast<const void*>(&out)) out = in;
  }
};

template <typename _Val> using CollectionFromPartialSumToRowBlockConverter = ConverterImpl<_Val>;
template <typename _Val> using CollectionToRowBlockConverter = ConverterImpl<_Val>;
template <typename _Val> using CollectionFromRowBlockConverter = ConverterImpl<_Val>;
template <typename _Val> using MatrixFromRowBlockToAllConverter = ConverterImpl<_Val>;
template <typename _Val> using MatrixFromColBlockToAllConverter = ConverterImpl<_Val>;
template <typename _Val> using MatrixFromRowBlockConverter = ConverterImpl<_Val>;
template <typename _Val> using MatrixToRowBlockConverter = ConverterImpl<_Val>;
template <typename _Val> using DummyConverter = ConverterImpl<_Val>;

// ---------------------------------------------------------------------------------------------
// The whole pipeline as one resident plan (mcnla_solver_*): A stays in HBM, Qs is never formed.
template <typename _Val>
class Solver {
 public:
  Solver(const Parameters<_Val>& p, const mcnla_solver_config& cfg) : params_(p), cfg_(cfg) {
    check(mcnla_solver_create(&params_.c(), &cfg_, &h_), "solver");
  }
  explicit Solver(const Parameters<_Val>& p) : params_(p) {
    mcnla_solver_config_default(&cfg_);
    check(mcnla_solver_create(&params_.c(), &cfg_, &h_), "solver");
  }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;
  ~Solver() {
    if (h_) mcnla_solver_destroy(h_);
  }
  // Solves for A given as a device pointer (resident) or a host-mirrored container.
  mcnla_status solve(const void* dA, index_t lda, int dtype, int order, void* stream = nullptr) {
    const mcnla_status s = mcnla_solver_solve(h_, dA, lda, dtype, order, stream);
    if (s != MCNLA_ERR_NOT_CONVERGED) check(s, "solve");
    return s;
  }
  template <class _Matrix>
  mcnla_status operator()(const _Matrix& a) {
    return solve(a.storage().device(), a.pitch(), _Matrix::dtype(), _Matrix::order());
  }
  index_t iteration() const {
    int64_t it = 0;
    check(mcnla_solver_iteration(h_, &it), "iteration");
    return it;
  }
  void results(const double** S, const double** U, const double** V, const double** Qbar) const {
    check(mcnla_solver_results(h_, S, U, V, Qbar), "results");
  }
  mcnla_solver handle() const { return h_; }

 private:
  const Parameters<_Val>& params_;
  mcnla_solver_config cfg_{};
  mcnla_solver h_ = nullptr;
};

}  // namespace isvd
}  // namespace mcnla

#endif  // MCNLA_AMD_ISVD_HPP_
