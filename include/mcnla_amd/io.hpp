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
#include <utility>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>

#include "isvd.hpp"

namespace mcnla {

// A half-open index range [begin, end) (the reference's IdxRange, core/utility/index_range.hpp),
// as the block loaders take it.
struct IdxRange {
  index_t begin = 0, end = 0;
  index_t len() const { return end - begin; }
};

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

// The size as the matrix of trans (io::loadMatrixMarketSize<Trans::NORMAL/TRANS>,
// matrix_market/dense_load.hpp): TRANS gives (n, m).
template <Trans trans>
void loadMatrixMarketSize(index_t& nrow, index_t& ncol, const char* file) {
  loadMatrixMarketSize(nrow, ncol, file);
  if (trans == Trans::TRANS) std::swap(nrow, ncol);
}

// Columns [cols.begin, cols.end) of a column-major MatrixMarket file
// (loadMatrixMarketColBlock, matrix_market/dense_load_block.hpp:62-103): the rank's column block
// of isvd_dense_cb.  An empty block is allocated m x len.
template <typename T, bool ROW>
void loadMatrixMarketColBlock(matrix::DenseMatrix<T, ROW>& block, const char* file,
                              const IdxRange& cols) {
  auto fin = open_in(file);
  skip_comments(fin);
  index_t m = 0, n = 0;
  fin >> m >> n;
  if (cols.begin < 0 || cols.end > n || cols.len() < 0)
    throw Error(MCNLA_ERR_INVALID, "column range outside the MatrixMarket matrix");
  if (block.isEmpty()) block.reconstruct(m, cols.len());
  if (block.nrow() != m || block.ncol() != cols.len())
    throw Error(MCNLA_ERR_INVALID, "column block size mismatch");
  double v;
  for (index_t e = 0; e < cols.begin * m; ++e) fin >> v;
  for (index_t j = 0; j < cols.len(); ++j)
    for (index_t i = 0; i < m; ++i) {
      fin >> v;
      block(i, j) = static_cast<T>(v);
    }
  if (!fin) throw Error(MCNLA_ERR_INVALID, std::string("truncated MatrixMarket file ") + file);
}

// Rows [rows.begin, rows.end) of a column-major MatrixMarket file
// (loadMatrixMarketRowBlock, dense_load_block.hpp:127-179): the rank's row block of
// isvd_dense_rb.  Each column's values outside the range are skipped (m - end of them after the
// block; the reference skips n - end, which agrees for square matrices).
template <typename T, bool ROW>
void loadMatrixMarketRowBlock(matrix::DenseMatrix<T, ROW>& block, const char* file,
                              const IdxRange& rows) {
  auto fin = open_in(file);
  skip_comments(fin);
  index_t m = 0, n = 0;
  fin >> m >> n;
  if (rows.begin < 0 || rows.end > m || rows.len() < 0)
    throw Error(MCNLA_ERR_INVALID, "row range outside the MatrixMarket matrix");
  if (block.isEmpty()) block.reconstruct(rows.len(), n);
  if (block.nrow() != rows.len() || block.ncol() != n)
    throw Error(MCNLA_ERR_INVALID, "row block size mismatch");
  double v;
  for (index_t j = 0; j < n; ++j) {
    for (index_t i = 0; i < rows.begin; ++i) fin >> v;
    for (index_t i = 0; i < rows.len(); ++i) {
      fin >> v;
      block(i, j) = static_cast<T>(v);
    }
    for (index_t i = rows.end; i < m; ++i) fin >> v;
  }
  if (!fin) throw Error(MCNLA_ERR_INVALID, std::string("truncated MatrixMarket file ") + file);
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

// The size as the matrix of trans (io::loadBinarySize<Trans::NORMAL/TRANS>, binary/dense_load.hpp):
// NORMAL reads the storage dims as (m, n) of a column-major matrix, TRANS as (n, m).
template <Trans trans>
void loadBinarySize(index_t& nrow, index_t& ncol, const char* file) {
  loadBinarySize<trans == Trans::TRANS>(nrow, ncol, file);
}

namespace detail {
// Opens a dense binary file of value type T; returns its storage dims.
template <typename T>
std::ifstream open_binary(const char* file, int64_t& d0, int64_t& d1) {
  std::ifstream in(file, std::ios::binary);
  if (!in) throw Error(MCNLA_ERR_INVALID, std::string("cannot open ") + file);
  if (read_header(in) != val_code<T>()) throw Error(MCNLA_ERR_INVALID, "binary value type mismatch");
  int64_t nd = 0;
  in.read(reinterpret_cast<char*>(&nd), 8);
  in.read(reinterpret_cast<char*>(&d0), 8);
  in.read(reinterpret_cast<char*>(&d1), 8);
  if (!in || nd != 2) throw Error(MCNLA_ERR_INVALID, "binary file is not a matrix");
  return in;
}
}  // namespace detail

// Storage lines [range.begin, range.end) of a dense binary file: the columns of a column-major
// matrix (loadBinaryColBlock, binary/dense_load_block.hpp:68-113) or the rows of a row-major one
// (loadBinaryRowBlock, :122-142, which reads the transpose's column block).  Seeks past the lines
// before the block, so a rank reads only its own bytes.
template <typename T, bool ROW>
void loadBinaryLines(matrix::DenseMatrix<T, ROW>& block, const char* file, const IdxRange& range) {
  int64_t d0 = 0, d1 = 0;
  auto in = detail::open_binary<T>(file, d0, d1);
  if (range.begin < 0 || range.end > d1 || range.len() < 0)
    throw Error(MCNLA_ERR_INVALID, "block range outside the binary matrix");
  const index_t m = ROW ? range.len() : d0, n = ROW ? d0 : range.len();
  if (block.isEmpty()) block.reconstruct(m, n);
  if (block.nrow() != m || block.ncol() != n) throw Error(MCNLA_ERR_INVALID, "block size mismatch");
  in.seekg(static_cast<std::streamoff>(d0 * range.begin * static_cast<int64_t>(sizeof(T))),
           std::ios_base::cur);
  in.read(reinterpret_cast<char*>(block.valPtr()),
          static_cast<std::streamsize>(d0 * range.len() * static_cast<int64_t>(sizeof(T))));
  if (!in) throw Error(MCNLA_ERR_INVALID, std::string("truncated binary file ") + file);
}

template <typename T>
void loadBinaryColBlock(matrix::DenseMatrix<T, false>& block, const char* file,
                        const IdxRange& cols) {
  loadBinaryLines(block, file, cols);
}

template <typename T>
void loadBinaryRowBlock(matrix::DenseMatrix<T, true>& block, const char* file,
                        const IdxRange& rows) {
  loadBinaryLines(block, file, rows);
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
