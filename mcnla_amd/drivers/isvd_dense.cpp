// Drop-in iSVD drivers for dense matrices on one MI355X: isvd_dense_cb (column-block layout,
// src/isvd_dense_cb.cpp) and, built with -DROW_BLOCK, isvd_dense_rb (row-block layout,
// src/isvd_dense_rb.cpp).  Same positional command line and the same outputs:
//
//   isvd_dense_cb <A-file> <S-file> <U-file> <V-file> [#sketch] [rank] [over-sampling-rank]
//                 [tolerance] [maxiter] [--seed S] [--omega philox|reference] [--threads T]
//
// The stage types are chosen at compile time with -DSTYPE=... -DOTYPE=... -DITYPE=... -DFTYPE=...
// exactly as in the reference (src/isvd_dense_cb.cpp:20-34); FILETYPE=Binary reads the MCNLA
// binary format instead of MatrixMarket.  The reference seeds Omega with rand() after
// srand(time(NULL)) (src/isvd_dense_cb.cpp:85,140); --seed makes a run reproducible.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <string>

#include "../../include/mcnla_amd/io.hpp"
#include "../../include/mcnla_amd/isvd.hpp"

#ifndef STYPE
#ifdef ROW_BLOCK
#define STYPE RowBlockGaussianProjectionSketcher
#else
#define STYPE ColBlockGaussianProjectionSketcher
#endif
#endif
#ifndef OTYPE
#define OTYPE RowBlockGramianOrthogonalizer
#endif
#ifndef ITYPE
#define ITYPE RowBlockWenYinIntegrator
#endif
#ifndef FTYPE
#ifdef ROW_BLOCK
#define FTYPE RowBlockGramianFormer
#else
#define FTYPE ColBlockGramianFormer
#endif
#endif
#ifdef NJOBV
#define JOBV false
#else
#define JOBV true
#endif

#ifdef ROW_BLOCK
using AMatrix = mcnla::DenseMatrixRowMajor<double>;
static const char* kLayout = "row-block";
#else
using AMatrix = mcnla::DenseMatrixColMajor<double>;
static const char* kLayout = "column-block";
#endif

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ||A - U S V'||_F / ||A||_F (src/isvd_dense_cb.cpp:308-332), on the host.
template <class UM>
double frerr(const AMatrix& a, const UM& u, const mcnla::DenseMatrixRowMajor<double>& v,
             const mcnla::DenseVector<double>& s) {
  const mcnla::index_t m = a.nrow(), n = a.ncol(), k = s.len();
  double num = 0.0, den = 0.0;
  for (mcnla::index_t i = 0; i < m; ++i)
    for (mcnla::index_t j = 0; j < n; ++j) {
      double usv = 0.0;
      for (mcnla::index_t e = 0; e < k; ++e) usv += u(i, e) * s(e) * v(j, e);
      const double x = a(i, j);
      num += (x - usv) * (x - usv);
      den += x * x;
    }
  return std::sqrt(num) / std::sqrt(den);
}

void usage(const char* prog) {
  std::cout << "Usage: " << prog
            << " <A-mtx-file> <S-mtx-file> <U-mtx-file> <V-mtx-file>"
               " [#sketch] [rank] [over-sampling-rank] [tolerance] [maxiter]"
               " [--seed S] [--omega philox|reference] [--threads T]"
            << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  bool have_seed = false;
  uint64_t seed = 0;
  int omega_mode = MCNLA_OMEGA_PHILOX, threads = 1;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--seed" && i + 1 < argc) {
      seed = std::strtoull(argv[++i], nullptr, 10);
      have_seed = true;
    } else if (a == "--omega" && i + 1 < argc) {
      std::string v = argv[++i];
      omega_mode = (v == "reference") ? MCNLA_OMEGA_REFERENCE : MCNLA_OMEGA_PHILOX;
    } else if (a == "--threads" && i + 1 < argc) {
      threads = std::atoi(argv[++i]);
    } else {
      pos.push_back(a);
    }
  }
  std::cout << "MCNLA (mcnla_amd) iSVD driver for DenseMatrix (" << kLayout << " version), "
            << mcnla_version() << std::endl << std::endl;
  if (pos.size() < 4) {
    usage(argv[0]);
    return 1;
  }
  try {
    // positional parameters, defaults of src/isvd_dense_cb.cpp:89-94
    size_t argi = 4;
    mcnla::index_t N = pos.size() > argi ? std::atof(pos[argi].c_str()) : 16;
    ++argi;
    mcnla::index_t k = pos.size() > argi ? std::atof(pos[argi].c_str()) : 20;
    ++argi;
    mcnla::index_t p = pos.size() > argi ? std::atof(pos[argi].c_str()) : 12;
    ++argi;
    double tol = pos.size() > argi ? std::atof(pos[argi].c_str()) : 1e-3;
    ++argi;
    mcnla::index_t maxiter = pos.size() > argi ? std::atof(pos[argi].c_str()) : 256;
    if (!have_seed) {
      std::srand(static_cast<unsigned>(std::time(nullptr)));
      seed = static_cast<uint64_t>(std::rand());
    }

    mcnla::index_t m = 0, n = 0;
#ifdef FILETYPE_BINARY
    mcnla::io::loadBinarySize<AMatrix::order() == MCNLA_ROW_MAJOR>(m, n, pos[0].c_str());
#else
    mcnla::io::loadMatrixMarketSize(m, n, pos[0].c_str());
#endif
    std::cout << "m = " << m << ", n = " << n << ", k = " << k << ", p = " << p << ", N = " << N
              << ", tol = " << tol << ", maxiter = " << maxiter << ", seed = " << seed << std::endl;
    std::cout << "1 GPU (one MI355X), " << (omega_mode == MCNLA_OMEGA_REFERENCE ? "reference" : "philox")
              << " Omega" << std::endl << std::endl;

    mcnla::isvd::Parameters<double> parameters(0, mcnla::kCommWorld);
    parameters.setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N);
    parameters.sync();

    mcnla::isvd::STYPE<double> sketcher(parameters);
    mcnla::isvd::OTYPE<double> orthogonalizer(parameters);
    mcnla::isvd::ITYPE<double> integrator(parameters);
    mcnla::isvd::FTYPE<double, JOBV> former(parameters);
#ifndef ROW_BLOCK
    mcnla::isvd::CollectionFromPartialSumToRowBlockConverter<double> so_converter(parameters);
#endif
    mcnla::isvd::MatrixFromRowBlockToAllConverter<double> if_converter(parameters);

    sketcher.setSeed(static_cast<mcnla::index_t>(seed));
    sketcher.setOmega(omega_mode, threads);
    integrator.setMaxIteration(maxiter).setTolerance(tol);
    sketcher.initialize();
    orthogonalizer.initialize();
    integrator.initialize();
    former.initialize();

    std::cout << "Uses " << sketcher << "." << std::endl;
    std::cout << "Uses " << orthogonalizer << "." << std::endl;
    std::cout << "Uses " << integrator << "." << std::endl;
    std::cout << "Uses " << former << "." << std::endl << std::endl;

    std::cout << "Reading data from " << pos[0] << "." << std::endl;
    AMatrix a;
    const double t_read = now();
#ifdef FILETYPE_BINARY
    mcnla::io::loadBinary(a, pos[0].c_str());
#else
    mcnla::io::loadMatrixMarket(a, pos[0].c_str());
#endif
    std::cout << "Done! (" << std::fixed << std::setprecision(6) << now() - t_read << " seconds)"
              << std::endl << std::endl;

    auto collection_q = parameters.createCollectionQj();
    auto matrix_qbar = parameters.createMatrixQbar();

    std::cout << "Start iSVD." << std::endl << std::endl;
    std::cout << "Sketching ............................. " << std::flush;
    sketcher(a, collection_q);
    std::cout << "Done!" << std::endl;
#ifndef ROW_BLOCK
    std::cout << "Scattering ............................ " << std::flush;
    so_converter(collection_q, collection_q);
    std::cout << "Done!" << std::endl;
#endif
    std::cout << "Orthogonalization ..................... " << std::flush;
    orthogonalizer(collection_q);
    std::cout << "Done!" << std::endl;
    std::cout << "Integration ........................... " << std::flush;
    integrator(collection_q, matrix_qbar);
    std::cout << "Done!" << std::endl;
    std::cout << "Gathering ............................. " << std::flush;
    if_converter(matrix_qbar, matrix_qbar);
    std::cout << "Done!" << std::endl;
    std::cout << "Forming ............................... " << std::flush;
    former(a, matrix_qbar);
    std::cout << "Done!" << std::endl << std::endl;

    const auto& s = former.vectorS();
    const auto& u = former.matrixU();
    const double ts = sketcher.time(), to = orthogonalizer.time(), ti = integrator.time(),
                 tf = former.time();
    std::cout << "Average total computing time:   " << ts + to + ti + tf << " seconds." << std::endl;
    std::cout << "Average sketching time:         " << ts << " seconds." << std::endl;
    std::cout << "Average orthogonalizing time:   " << to << " seconds." << std::endl;
    std::cout << "Average integrating time:       " << ti << " seconds." << std::endl;
    std::cout << "Average forming time:           " << tf << " seconds." << std::endl;
    std::cout << std::endl << "Iteration = " << integrator.iteration() << std::endl;
    if (JOBV) {
      std::cout << "Error     = " << std::scientific << std::setprecision(6)
                << frerr(a, u, former.matrixV(), s) << std::endl << std::endl;
      std::cout << "Error := norm(A - Uk Sk Vk')_F / norm(A)_F" << std::endl;
    }
    std::cout << std::endl;
    std::cout << "Write S into " << pos[1] << "." << std::endl;
    mcnla::io::saveMatrixMarket(s, pos[1].c_str());
    std::cout << "Write U into " << pos[2] << "." << std::endl;
    mcnla::io::saveMatrixMarket(u, pos[2].c_str());
    std::cout << "Write V into " << pos[3] << "." << std::endl;
    if (JOBV) mcnla::io::saveMatrixMarket(former.matrixV(), pos[3].c_str());
    else mcnla::io::saveMatrixMarket(mcnla::DenseMatrixRowMajor<double>(), pos[3].c_str());
    std::cout << std::endl;
  } catch (const mcnla::Error& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
