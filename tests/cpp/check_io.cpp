// Exercises the host file formats (include/mcnla_amd/io.hpp) for the CPU tests; no device use.
//   save <m> <n> <C|F> <file> <comment>   writes a(i, j) = 1000 i + j (MCNLA binary)
//   size <file> <N|T>                     prints loadBinarySize<Trans>
//   mtxsize <file> <N|T>                  prints loadMatrixMarketSize<Trans>
//   colblock <file> <j0> <j1>             loadBinaryColBlock, prints the block column-major
//   rowblock <file> <i0> <i1>             loadBinaryRowBlock, prints the block row-major
//   mtxcol <file> <j0> <j1>               loadMatrixMarketColBlock (column-major block)
//   mtxrow <file> <i0> <i1>               loadMatrixMarketRowBlock (row-major block)
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "../../include/mcnla_amd/io.hpp"

using mcnla::index_t;

template <typename M>
void print(const M& a) {
  std::cout << a.nrow() << " " << a.ncol() << "\n";
  std::cout.precision(17);
  for (index_t i = 0; i < a.nrow(); ++i)
    for (index_t j = 0; j < a.ncol(); ++j) std::cout << a(i, j) << (j + 1 < a.ncol() ? " " : "\n");
}

int main(int argc, char** argv) {
  if (argc < 3) return 1;
  const std::string cmd = argv[1];
  try {
    if (cmd == "save" && argc >= 7) {
      const index_t m = std::atoll(argv[2]), n = std::atoll(argv[3]);
      if (argv[4][0] == 'C') {
        mcnla::DenseMatrixRowMajor<double> a(m, n);
        for (index_t i = 0; i < m; ++i)
          for (index_t j = 0; j < n; ++j) a(i, j) = 1000.0 * i + j;
        mcnla::io::saveBinary(a, argv[5], argv[6]);
      } else {
        mcnla::DenseMatrixColMajor<double> a(m, n);
        for (index_t i = 0; i < m; ++i)
          for (index_t j = 0; j < n; ++j) a(i, j) = 1000.0 * i + j;
        mcnla::io::saveBinary(a, argv[5], argv[6]);
      }
    } else if ((cmd == "size" || cmd == "mtxsize") && argc >= 4) {
      index_t m = 0, n = 0;
      const bool t = argv[3][0] == 'T';
      if (cmd == "size") {
        if (t) mcnla::io::loadBinarySize<mcnla::Trans::TRANS>(m, n, argv[2]);
        else mcnla::io::loadBinarySize<mcnla::Trans::NORMAL>(m, n, argv[2]);
      } else {
        if (t) mcnla::io::loadMatrixMarketSize<mcnla::Trans::TRANS>(m, n, argv[2]);
        else mcnla::io::loadMatrixMarketSize<mcnla::Trans::NORMAL>(m, n, argv[2]);
      }
      std::cout << m << " " << n << "\n";
    } else if (argc >= 5) {
      const mcnla::IdxRange r{std::atoll(argv[3]), std::atoll(argv[4])};
      if (cmd == "colblock") {
        mcnla::DenseMatrixColMajor<double> b;
        mcnla::io::loadBinaryColBlock(b, argv[2], r);
        print(b);
      } else if (cmd == "rowblock") {
        mcnla::DenseMatrixRowMajor<double> b;
        mcnla::io::loadBinaryRowBlock(b, argv[2], r);
        print(b);
      } else if (cmd == "mtxcol") {
        mcnla::DenseMatrixColMajor<double> b;
        mcnla::io::loadMatrixMarketColBlock(b, argv[2], r);
        print(b);
      } else if (cmd == "mtxrow") {
        mcnla::DenseMatrixRowMajor<double> b;
        mcnla::io::loadMatrixMarketRowBlock(b, argv[2], r);
        print(b);
      } else {
        return 1;
      }
    } else {
      return 1;
    }
  } catch (const mcnla::Error& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
