// Caller-owned device-memory views used by the drop-in entry points.
//
// They stand in for the cusp views the reference's signatures take
// (cusp::array1d<T, device_memory>::view / const_view,
// cusp::array2d<real, device_memory>::view / const_view, and
// thrust::device_ptr<void> for temp; see
// split/include/split/device/kmeans/cluster.cuh:31-37) and expose the members
// the reference's callers read: num_rows, num_cols, num_entries, values,
// row(i), size(), begin()/end(), get().  A view never owns memory; the
// pointer is a device address.  Dense layout only (pitch == num_cols), which
// is what the reference's callers assume (view_util.cuh:16-19).
#pragma once

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "split/detail/internal.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

template <typename T>
class device_ptr {
 public:
  device_ptr() = default;
  explicit device_ptr(T* p) : p_(p) {}
  template <typename U, typename = std::enable_if_t<std::is_convertible<U*, T*>::value>>
  device_ptr(const device_ptr<U>& o) : p_(o.get()) {}
  T* get() const { return p_; }

 private:
  T* p_ = nullptr;
};

template <typename T>
device_ptr<T> device_pointer_cast(T* p) { return device_ptr<T>(p); }

template <typename T>
class array1d_view {
 public:
  array1d_view() = default;
  array1d_view(T* data, std::int64_t size) : data_(data), size_(size) {}
  template <typename U, typename = std::enable_if_t<std::is_convertible<U*, T*>::value>>
  array1d_view(const array1d_view<U>& o) : data_(o.data()), size_(o.size()) {}

  T* data() const { return data_; }
  std::int64_t size() const { return size_; }
  T* begin() const { return data_; }
  T* end() const { return data_ + size_; }
  array1d_view subarray(std::int64_t offset, std::int64_t count) const {
    return array1d_view(data_ + offset, count);
  }

 private:
  T* data_ = nullptr;
  std::int64_t size_ = 0;
};

template <typename T>
class array2d_view {
 public:
  array2d_view() = default;
  array2d_view(std::int64_t rows, std::int64_t cols, T* data)
      : num_rows(rows), num_cols(cols), num_entries(rows * cols),
        values(data, rows * cols) {}
  template <typename U, typename = std::enable_if_t<std::is_convertible<U*, T*>::value>>
  array2d_view(const array2d_view<U>& o)
      : num_rows(o.num_rows), num_cols(o.num_cols),
        num_entries(o.num_entries), values(o.values) {}

  array1d_view<T> row(std::int64_t i) const {
    return values.subarray(i * num_cols, num_cols);
  }

  std::int64_t num_rows = 0;
  std::int64_t num_cols = 0;
  std::int64_t num_entries = 0;
  array1d_view<T> values;
};

// The names the reference's signatures use, as aliases.
template <typename T> using view1d = array1d_view<T>;
template <typename T> using const_view1d = array1d_view<const T>;
template <typename T> using view2d = array2d_view<T>;
template <typename T> using const_view2d = array2d_view<const T>;

template <typename T>
array2d_view<T> make_array2d_view(std::int64_t rows, std::int64_t cols, T* data) {
  return array2d_view<T>(rows, cols, data);
}
template <typename T>
array1d_view<T> make_array1d_view(T* data, std::int64_t size) {
  return array1d_view<T>(data, size);
}

SPLIT_DEVICE_NAMESPACE_END
