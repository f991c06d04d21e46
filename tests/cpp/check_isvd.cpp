// The reference's stage tests (check/isvd/**) restated against the mcnla_amd C++ API, on the GPU.
// Usage: check_isvd <golden-dir>.  Prints one line per test; exit status = number of failures.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/mcnla_amd/io.hpp"
#include "../../include/mcnla_amd/isvd.hpp"

using mcnla::index_t;
using Coll = mcnla::DenseMatrixCollectionColBlockRowMajor<double>;
using RowM = mcnla::DenseMatrixRowMajor<double>;
using ColM = mcnla::DenseMatrixColMajor<double>;

static std::string g_dir;
static int g_fail = 0;

static std::string path(const char* f) { return g_dir + "/" + f; }

static void report(const char* name, bool ok, const std::string& detail) {
  std::printf("%s %s %s\n", ok ? "PASS" : "FAIL", name, detail.c_str());
  std::fflush(stdout);
  if (!ok) ++g_fail;
}

// max |P1 - P2| over the lower triangle of the projectors of the first k columns
template <class M1, class M2>
static double projector_diff(const M1& a, const M2& b, index_t m, index_t k) {
  double worst = 0.0;
  for (index_t i = 0; i < m; ++i)
    for (index_t j = 0; j <= i; ++j) {
      double x = 0.0, y = 0.0;
      for (index_t e = 0; e < k; ++e) {
        x += a(i, e) * a(j, e);
        y += b(i, e) * b(j, e);
      }
      worst = std::fmax(worst, std::fabs(x - y));
    }
  return worst;
}

// check/isvd/integrator/*.cpp
template <template <typename> class Integrator>
static void integrator_test(const char* name, const char* expected, double tol, index_t iters) {
  Coll qi;
  RowM qbar_true;
  mcnla::io::loadMatrixMarket(qi, path("qi.mtx").c_str());
  mcnla::io::loadMatrixMarket(qbar_true, path(expected).c_str());
  const index_t m = qi.nrow(), k = qi.ncol(), N = qi.nmat();
  mcnla::isvd::Parameters<double> parameters(0, mcnla::kCommWorld);
  parameters.setSize(m, k).setRank(k).setOverRank(0).setNumSketchEach(N);
  parameters.sync();
  Integrator<double> integrator(parameters, 256, tol);
  integrator.initialize();
  auto qbar = parameters.createMatrixQbar();
  integrator(qi, qbar);
  double err = 0.0;
  for (index_t i = 0; i < m; ++i)
    for (index_t j = 0; j < k; ++j) err = std::fmax(err, std::fabs(qbar(i, j) - qbar_true(i, j)));
  char buf[160];
  std::snprintf(buf, sizeof(buf), "iteration %ld (expect %ld), max |Qbar - Qbar_true| %.2e",
                (long)integrator.iteration(), (long)iters, err);
  report(name, integrator.iteration() == iters && err < 1e-8, buf);
}

// check/isvd/orthogonalizer/*gramian_orthogonalizer.cpp
template <template <typename> class Orthogonalizer>
static void orthogonalizer_test(const char* name) {
  Coll yi, qi_true;
  mcnla::io::loadMatrixMarket(yi, path("yi.mtx").c_str());
  mcnla::io::loadMatrixMarket(qi_true, path("qi.mtx").c_str());
  const index_t m = yi.nrow(), k = yi.ncol(), N = yi.nmat();
  mcnla::isvd::Parameters<double> parameters(0, mcnla::kCommWorld);
  parameters.setSize(m, k).setRank(k).setOverRank(0).setNumSketchEach(N);
  parameters.sync();
  Orthogonalizer<double> orthogonalizer(parameters);
  orthogonalizer.initialize();
  orthogonalizer(yi);
  double err = 0.0;
  for (index_t s = 0; s < N; ++s) {
    auto a = [&](index_t i, index_t j) { return yi(i, j, s); };
    auto b = [&](index_t i, index_t j) { return qi_true(i, j, s); };
    err = std::fmax(err, projector_diff(a, b, m, k));
  }
  char buf[96];
  std::snprintf(buf, sizeof(buf), "max projector diff %.2e", err);
  report(name, err < 1e-8, buf);
}

// check/isvd/former/*.cpp
template <class Former, class AM>
static void former_test(const char* name) {
  AM a;
  RowM q;
  ColM u_true;
  mcnla::io::loadMatrixMarket(a, path("a.mtx").c_str());
  mcnla::io::loadMatrixMarket(q, path("q.mtx").c_str());
  mcnla::io::loadMatrixMarket(u_true, path("u.mtx").c_str());
  const index_t m = a.nrow(), n = a.ncol(), k = q.ncol() / 2, p = q.ncol() - k;
  mcnla::isvd::Parameters<double> parameters(0, mcnla::kCommWorld);
  parameters.setSize(m, n).setRank(k).setOverRank(p).setNumSketchEach(1);
  parameters.sync();
  Former former(parameters);
  former.initialize();
  former(a, q);
  const double err = projector_diff(former.matrixU(), u_true, m, k);
  char buf[96];
  std::snprintf(buf, sizeof(buf), "max projector diff %.2e", err);
  report(name, err < 1e-8, buf);
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("usage: %s <golden-dir>\n", argv[0]);
    return 100;
  }
  g_dir = argv[1];
  try {
    namespace I = mcnla::isvd;
    integrator_test<I::KolmogorovNagumoIntegrator>("KolmogorovNagumoIntegrator", "qb_kn.mtx", 1e-4, 81);
    integrator_test<I::RowBlockKolmogorovNagumoIntegrator>("RowBlockKolmogorovNagumoIntegrator", "qb_kn.mtx", 1e-4, 81);
    integrator_test<I::RowBlockGramianKolmogorovNagumoIntegrator>("RowBlockGramianKolmogorovNagumoIntegrator", "qb_kn.mtx", 1e-4, 81);
    integrator_test<I::RowBlockWenYinIntegrator>("RowBlockWenYinIntegrator", "qb_wy.mtx", 1e-3, 21);
    integrator_test<I::RowBlockGramianWenYinIntegrator>("RowBlockGramianWenYinIntegrator", "qb_wy.mtx", 1e-3, 21);
    orthogonalizer_test<I::GramianOrthogonalizer>("GramianOrthogonalizer");
    orthogonalizer_test<I::RowBlockGramianOrthogonalizer>("RowBlockGramianOrthogonalizer");
    former_test<I::SvdFormer<double, true>, RowM>("SvdFormer");
    former_test<I::GramianFormer<double, true>, ColM>("GramianFormer");
    former_test<I::ColBlockGramianFormer<double, true>, ColM>("ColBlockGramianFormer");
    former_test<I::RowBlockGramianFormer<double, true>, RowM>("RowBlockGramianFormer");
  } catch (const mcnla::Error& e) {
    std::printf("FAIL exception %s\n", e.what());
    return 99;
  }
  return g_fail;
}
