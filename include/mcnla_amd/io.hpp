// Host-side file formats of the reference, for the drop-in drivers (host code, no device use):
//   MatrixMarket dense "array" text: header lines starting with '%', then "m n" (or "m n l" for
//     a collection), then the values in column-major order, one matrix after another
//     (include/mcnla/core/io/matrix_market/dense_load.hpp:87-121, :198-235; dense_save.hpp);
//   MCNLA binary: a 128-byte text header -- "##MCNLABinaryFileFormat" padded to 31 + '\n',
//     "# <comment>" padded to 63 + '\n', storage "DEN" in 8 bytes, value type "R8"/"R4" in 8,
//     8 blank (index type, dense), 7 blank + '\n' -- then int64 ndim, int64 dims in storage order
//     (dim0 = the leading dimension), then the raw values in storage order
//     (include/mcnla/core/io/binary/def.hpp:77-154, binary/dense_save.hpp:81-90).
#ifndef MCNLA_AMD_IO_HPP_
#define MCNLA_AMD_IO_HPP_

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>

#include "isvd.hpp"

namespace mcnla {
namespace io {

inline std::ifstream open_in(const char* file) {
  std::ifstream fin(file);
  if (!fin) throw Error(MCNLA_ERR_INVALID, std::string("cannot open ") + file);
  return fin;
}

inline void skip_comments(std::istream& fin) {
  while (fin.peek() == '%') fin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
}

// ---------------------------------------------------------------------------------------------
// MatrixMarket
inline void loadMatrixMarketSize(index_t& nrow, index_t& ncol, const char* file) {
  auto fin = open_in(file);
  skip_comments(fin);
  fin >> nrow >> ncol;
  if (!fin) throw Error(MCNLA_ERR_INVALID, std::string("bad MatrixMarket size line in ") + file);
}

template <typename T, bool ROW>
void loadMatrixMarket(matrix::DenseMatrix<T, ROW>& a, const char* file) {
  auto fin = open_in(file);
  skip_comments(fin);
  index_t m = 0, n = 0;
  fin >> m >> n;
  if (a.isEmpty()) a.reconstruct(m, n);
  if (a.nrow() != m || a.ncol() != n) throw Error(MCNLA_ERR_INVALID, "MatrixMarket size mismatch");
  T* p = a.valPtr();
  for (index_t j = 0; j < n; ++j)
    for (index_t i = 0; i < m; ++i) {
      double v;
      fin >> v;
      p[ROW ? i * n + j : i + j * m] = static_cast<T>(v);
    }
  if (!fin) throw Error(MCNLA_ERR_INVALID, std::string("truncated MatrixMarket file ") + file);
}

template <typename T>
void loadMatrixMarket(matrix::DenseMatrixCollectionColBlockRowMajor<T>& c, const char* file) {
  auto fin = open_in(file);
  skip_comments(fin);
  index_t m = 0, n = 0, l = 0;
  fin >> m >> n >> l;
  if (c.isEmpty()) c.reconstruct(m, n, l);
  T* p = c.unfold().valPtr();
  const index_t ld = n * l;
  for (index_t k = 0; k < l; ++k)
    for (index_t j = 0; j < n; ++j)
      for (index_t i = 0; i < m; ++i) {
        double v;
        fin >> v;
        p[i * ld + k * n + j] = static_cast<T>(v);
      }
  if (!fin) throw Error(MCNLA_ERR_INVALID, std::string("truncated MatrixMarket file ") + file);
}

template <typename T, bool ROW>
void saveMatrixMarket(const matrix::DenseMatrix<T, ROW>& a, const char* file) {
  std::ofstream fout(file);
  if (!fout) throw Error(MCNLA_ERR_INVALID, std::string("cannot write ") + file);
  fout << "%%MatrixMarket matrix array real general\n" << a.nrow() << " " << a.ncol() << "\n";
  fout << std::setprecision(16);
  const T* p = a.valPtr();
  for (index_t j = 0; j < a.ncol(); ++j)
    for (index_t i = 0; i < a.nrow(); ++i)
      fout << p[ROW ? i * a.ncol() + j : i + j * a.nrow()] << "\n";
}

template <typename T>
void saveMatrixMarket(const matrix::DenseVector<T>& v, const char* file) {
  std::ofstream fout(file);
  if (!fout) throw Error(MCNLA_ERR_INVALID, std::string("cannot write ") + file);
  fout << "%%MatrixMarket matrix array real general\n" << v.len() << " 1\n" << std::setprecision(16);
  for (index_t i = 0; i < v.len(); ++i) fout << v(i) << "\n";
}

// ---------------------------------------------------------------------------------------------
// MCNLA binary
namespace detail {
constexpr int kHeaderLen = 32, kCommentLen = 64, kFieldLen = 8;

template <typename T>
const char* val_code() {
  return sizeof(T) == 8 ? "R8" : "R4";
}

template <typename T>
void write_header(std::ostream& out, const char* comment) {
  std::ostringstream h;
  h << std::left << std::setw(kHeaderLen - 1) << "##MCNLABinaryFileFormat" << '\n';
  h << "# " << std::left << std::setw(kCommentLen - 3) << comment << '\n';
  h << std::left << std::setw(kFieldLen) << "DEN" << std::setw(kFieldLen) << val_code<T>()
    << std::setw(kFieldLen) << "" << std::setw(kFieldLen - 1) << "" << '\n';
  const std::string s = h.str();
  if (s.size() != 128) throw Error(MCNLA_ERR_INVALID, "binary comment too long");
  out.write(s.data(), static_cast<std::streamsize>(s.size()));
}

// Returns the value-type code of the file ("R8" / "R4") after checking the storage field.
inline std::string read_header(std::istream& in) {
  char buf[128];
  in.read(buf, 128);
  if (!in || std::strncmp(buf, "##MCNLABinaryFileFormat", 23) != 0)
    throw Error(MCNLA_ERR_INVALID, "not an MCNLA binary file");
  if (std::strncmp(buf + kHeaderLen + kCommentLen, "DEN", 3) != 0)
    throw Error(MCNLA_ERR_INVALID, "only dense MCNLA binary files are supported");
  return std::string(buf + kHeaderLen + kCommentLen + kFieldLen, 2);
}
}  // namespace detail

// Matrix size (m, n) of a dense binary file that holds a column-major (ROW = false) or row-major
// (ROW = true) matrix.
template <bool ROW>
void loadBinarySize(index_t& nrow, index_t& ncol, const char* file) {
  std::ifstream in(file, std::ios::binary);
  if (!in) throw Error(MCNLA_ERR_INVALID, std::string("cannot open ") + file);
  detail::read_header(in);
  int64_t nd = 0, d0 = 0, d1 = 0;
  in.read(reinterpret_cast<char*>(&nd), 8);
  in.read(reinterpret_cast<char*>(&d0), 8);
  in.read(reinterpret_cast<char*>(&d1), 8);
  if (nd != 2) throw Error(MCNLA_ERR_INVALID, "binary file is not a matrix");
  nrow = ROW ? d1 : d0;
  ncol = ROW ? d0 : d1;
}

template <typename T, bool ROW>
void loadBinary(matrix::DenseMatrix<T, ROW>& a, const char* file) {
  std::ifstream in(file, std::ios::binary);
  if (!in) throw Error(MCNLA_ERR_INVALID, std::string("cannot open ") + file);
  const std::string code = detail::read_header(in);
  if (code != detail::val_code<T>()) throw Error(MCNLA_ERR_INVALID, "binary value type mismatch");
  int64_t nd = 0, d0 = 0, d1 = 0;
  in.read(reinterpret_cast<char*>(&nd), 8);
  in.read(reinterpret_cast<char*>(&d0), 8);
  in.read(reinterpret_cast<char*>(&d1), 8);
  if (nd != 2) throw Error(MCNLA_ERR_INVALID, "binary file is not a matrix");
  const index_t m = ROW ? d1 : d0, n = ROW ? d0 : d1;
  if (a.isEmpty()) a.reconstruct(m, n);
  if (a.nrow() != m || a.ncol() != n) throw Error(MCNLA_ERR_INVALID, "binary size mismatch");
  in.read(reinterpret_cast<char*>(a.valPtr()), static_cast<std::streamsize>(m * n * sizeof(T)));
  if (!in) throw Error(MCNLA_ERR_INVALID, std::string("truncated binary file ") + file);
}

template <typename T, bool ROW>
void saveBinary(const matrix::DenseMatrix<T, ROW>& a, const char* file, const char* comment = "") {
  std::ofstream out(file, std::ios::binary);
  if (!out) throw Error(MCNLA_ERR_INVALID, std::string("cannot write ") + file);
  detail::write_header<T>(out, comment);
  const int64_t nd = 2, d0 = ROW ? a.ncol() : a.nrow(), d1 = ROW ? a.nrow() : a.ncol();
  out.write(reinterpret_cast<const char*>(&nd), 8);
  out.write(reinterpret_cast<const char*>(&d0), 8);
  out.write(reinterpret_cast<const char*>(&d1), 8);
  out.write(reinterpret_cast<const char*>(a.valPtr()),
            static_cast<std::streamsize>(a.nrow() * a.ncol() * sizeof(T)));
}

}  // namespace io
}  // namespace mcnla

#endif  // MCNLA_AMD_IO_HPP_
