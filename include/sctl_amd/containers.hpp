// Minimal host containers with the parts of the sctl::Vector / sctl::Matrix
// interface that GenericKernel uses (vector.hpp:24-436, matrix.hpp:45-360):
// Dim(), ReInit(), SetZero(), begin(), operator[], row-major Matrix rows.
// For hosts that do not have SCTL; sctl::Vector works with the drop-in as is.
#ifndef SCTL_AMD_CONTAINERS_HPP_
#define SCTL_AMD_CONTAINERS_HPP_

#include <cstdint>
#include <algorithm>
#include <initializer_list>
#include <vector>

namespace sctl_amd {

template <class T> class Vector {
 public:
  Vector() = default;
  explicit Vector(int64_t n) : d_(n) {}
  Vector(std::initializer_list<T> v) : d_(v) {}
  explicit Vector(const std::vector<T>& v) : d_(v) {}
  int64_t Dim() const { return (int64_t)d_.size(); }
  void ReInit(int64_t n) { d_.resize(n); }  // keeps capacity, does not zero (vector.txx:102-130)
  void SetZero() { std::fill(d_.begin(), d_.end(), T(0)); }
  T* begin() { return d_.data(); }
  const T* begin() const { return d_.data(); }
  T& operator[](int64_t i) { return d_[i]; }
  const T& operator[](int64_t i) const { return d_[i]; }

 private:
  std::vector<T> d_;
};

template <class T> class Matrix {
 public:
  Matrix() = default;
  Matrix(int64_t r, int64_t c) : r_(r), c_(c), d_(r * c) {}
  int64_t Dim(int i) const { return i == 0 ? r_ : c_; }
  void ReInit(int64_t r, int64_t c) {
    r_ = r;
    c_ = c;
    d_.resize(r * c);
  }
  void SetZero() { std::fill(d_.begin(), d_.end(), T(0)); }
  T* begin() { return d_.data(); }
  const T* begin() const { return d_.data(); }
  T* operator[](int64_t i) { return d_.data() + i * c_; }
  const T* operator[](int64_t i) const { return d_.data() + i * c_; }
  T& operator()(int64_t i, int64_t j) { return d_[i * c_ + j]; }
  const T& operator()(int64_t i, int64_t j) const { return d_[i * c_ + j]; }

 private:
  int64_t r_ = 0, c_ = 0;
  std::vector<T> d_;
};

}  // namespace sctl_amd

#endif  // SCTL_AMD_CONTAINERS_HPP_
