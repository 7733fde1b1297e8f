// Distance types of the efanna API (reference: general/distance.hpp).  The
// MI355X library computes squared L2 on the device; these classes select it
// and keep a host compare() in the reference's AVX accumulation order (eight
// lane chains of fused multiply-adds, then a left-to-right sum) for callers
// that use them directly.
#ifndef EFANNA_DISTANCE_H_
#define EFANNA_DISTANCE_H_
#include <cmath>
#include <cstddef>

// Allocation alignment for row buffers handed to Matrix (reference:
// general/distance.hpp:10-20, used by the samples' memalign calls).  Rows
// here always have the round8 stride, so 32-byte alignment suits every build.
#ifndef KGRAPH_MATRIX_ALIGN
#define KGRAPH_MATRIX_ALIGN 32
#endif

namespace efanna {

template <typename T>
struct Candidate {
  size_t row_id;
  T distance;
  Candidate(const size_t r, const T d) : row_id(r), distance(d) {}
  bool operator>(const Candidate& o) const { return distance == o.distance ? row_id > o.row_id : distance > o.distance; }
  bool operator<(const Candidate& o) const { return distance == o.distance ? row_id < o.row_id : distance < o.distance; }
};

template <typename T>
class Distance {
 public:
  virtual T compare(const T* a, const T* b, size_t length) const = 0;
  virtual T norm(const T* a, size_t length) const = 0;
  virtual T dot(const T* a, const T* b, size_t length) const = 0;
  virtual ~Distance() {}
};

// Squared L2 (no square root) over rows padded to a multiple of 8.
template <typename T>
class L2DistanceAVX : public Distance<T> {
 public:
  T compare(const T* a, const T* b, size_t d) const {
    size_t D = (d + 7) & ~size_t(7), DR = D % 32, DD = D - DR;
    T acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t c = DD; c < D; c += 8)
      for (int j = 0; j < 8; ++j) {
        T t = a[c + j] - b[c + j];
        acc[j] = std::fma(t, t, acc[j]);
      }
    for (size_t i = 0; i < DD; i += 8)
      for (int j = 0; j < 8; ++j) {
        T t = a[i + j] - b[i + j];
        acc[j] = std::fma(t, t, acc[j]);
      }
    T s = acc[0];
    for (int j = 1; j < 8; ++j) s = s + acc[j];
    return s;
  }
  T norm(const T* a, size_t d) const {
    T s = 0;
    for (size_t i = 0; i < d; ++i) s += a[i] * a[i];
    return s;
  }
  T dot(const T* a, const T* b, size_t d) const {
    T s = 0;
    for (size_t i = 0; i < d; ++i) s += a[i] * b[i];
    return s;
  }
};
// The reference's other L2 variants differ only in host summation order; the
// library runs the same device kernel for all of them.
template <typename T>
class L2Distance : public L2DistanceAVX<T> {};
template <typename T>
class L2DistanceSSE : public L2DistanceAVX<T> {};
template <typename T>
class L2DistanceAVXr4 : public L2DistanceAVX<T> {};

}  // namespace efanna
#endif
