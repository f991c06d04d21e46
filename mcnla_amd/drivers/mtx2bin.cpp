// Converts a dense MatrixMarket file to the MCNLA binary format (src/mtx2bin.cpp), storing the
// matrix column-major (the layout isvd_dense_cb reads) or, with --row-major, row-major (the
// layout isvd_dense_rb reads).
#include <cstring>
#include <iostream>

#include "../../include/mcnla_amd/io.hpp"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cout << "Usage: " << argv[0] << " <input-mtx> <output-bin> [--row-major]" << std::endl;
    return 1;
  }
  const bool row = argc > 3 && std::strcmp(argv[3], "--row-major") == 0;
  try {
    if (row) {
      mcnla::DenseMatrixRowMajor<double> a;
      mcnla::io::loadMatrixMarket(a, argv[1]);
      mcnla::io::saveBinary(a, argv[2], "mtx2bin");
    } else {
      mcnla::DenseMatrixColMajor<double> a;
      mcnla::io::loadMatrixMarket(a, argv[1]);
      mcnla::io::saveBinary(a, argv[2], "mtx2bin");
    }
  } catch (const mcnla::Error& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
