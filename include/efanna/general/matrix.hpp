// Matrix view of the efanna API (reference: general/matrix.hpp:13-71): rows
// of `cols` values, row i at data + i * round8(cols) (the reference's AVX
// alignment); the view does not own the data.
#ifndef EFANNA_MATRIX_H
#define EFANNA_MATRIX_H
#include <cstddef>
#include <stdexcept>
#include <vector>

namespace efanna {
template <typename T>
class Matrix {
 public:
  Matrix(const size_t rows, const size_t cols, const void* data)
      : rows_(rows), cols_(cols), stride_((cols + 7) / 8 * 8), data_(reinterpret_cast<const T*>(data)) {}
  size_t get_cols() const { return cols_; }
  size_t get_rows() const { return rows_; }
  size_t get_stride() const { return stride_; }
  const T* data() const { return data_; }
  const T* get_row(const size_t index) const {
    if (index >= rows_) throw std::runtime_error("index out of range");
    return data_ + stride_ * index;
  }

 private:
  size_t rows_, cols_, stride_;
  const T* data_;
};
}  // namespace efanna
#endif
