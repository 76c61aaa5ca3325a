/** @file
 * Minimal host array view standing in for the blitz::Array views the
 * reference returns from advectee() / advector() (concurr/detail/
 * sharedmem.hpp:522-577): non-owning, arbitrary lower bounds, C-order
 * strides, element access with operator(), scalar assignment and
 * index-function fill in place of Blitz tensor expressions.
 */
#pragma once

#include <array>
#include <cstddef>
#include <stdexcept>

namespace libmpdataxx
{
  template <typename real_t, int n_dims>
  class arr_t
  {
    real_t *data_ = nullptr;              // element at the lower bounds
    std::array<int, n_dims> lbound_{}, extent_{};
    std::array<std::ptrdiff_t, n_dims> stride_{};

    public:

    arr_t() = default;
    arr_t(real_t *first, const std::array<int, n_dims> &lbound, const std::array<int, n_dims> &extent)
      : data_(first), lbound_(lbound), extent_(extent)
    {
      std::ptrdiff_t s = 1;
      for (int d = n_dims - 1; d >= 0; --d) { stride_[d] = s; s *= extent[d]; }
    }

    int lbound(int d) const { return lbound_[d]; }
    int ubound(int d) const { return lbound_[d] + extent_[d] - 1; }
    int extent(int d) const { return extent_[d]; }
    std::array<int, n_dims> base() const { return lbound_; }
    std::array<int, n_dims> shape() const { return extent_; }
    std::size_t size() const { std::size_t n = 1; for (int d = 0; d < n_dims; ++d) n *= extent_[d]; return n; }
    // like blitz::Array, a const view gives read-only elements
    real_t *dataFirst() { return data_; }
    const real_t *dataFirst() const { return data_; }
    std::ptrdiff_t stride(int d) const { return stride_[d]; }

    template <class... I>
    real_t &operator()(I... idx)
    {
      return data_[offset(idx...)];
    }

    template <class... I>
    const real_t &operator()(I... idx) const
    {
      return data_[offset(idx...)];
    }

    private:

    template <class... I>
    std::ptrdiff_t offset(I... idx) const
    {
      static_assert(sizeof...(I) == n_dims, "wrong number of indices");
      const int ix[n_dims] = {idx...};
      std::ptrdiff_t off = 0;
      for (int d = 0; d < n_dims; ++d) off += (ix[d] - lbound_[d]) * stride_[d];
      return off;
    }

    public:

    // blitz-like scalar assignment
    arr_t &operator=(real_t v)
    {
      for (std::size_t q = 0; q < size(); ++q) data_[q] = v;
      return *this;
    }

    // element-wise copy from a view of the same shape
    arr_t &assign(const arr_t &o)
    {
      if (o.shape() != shape()) throw std::runtime_error("libmpdata++: shape mismatch");
      for (std::size_t q = 0; q < size(); ++q) data_[q] = o.data_[q];
      return *this;
    }

    // fill from an index function f(i[, j[, k]]) over the view's own bounds
    template <class F>
    arr_t &fill(F f)
    {
      if constexpr (n_dims == 2)
      {
        for (int i = lbound(0); i <= ubound(0); ++i)
          for (int j = lbound(1); j <= ubound(1); ++j)
            (*this)(i, j) = f(i, j);
      }
      else
      {
        for (int i = lbound(0); i <= ubound(0); ++i)
          for (int j = lbound(1); j <= ubound(1); ++j)
            for (int k = lbound(2); k <= ubound(2); ++k)
              (*this)(i, j, k) = f(i, j, k);
      }
      return *this;
    }

    real_t min() const { real_t m = data_[0]; for (std::size_t q = 1; q < size(); ++q) m = data_[q] < m ? data_[q] : m; return m; }
    real_t max() const { real_t m = data_[0]; for (std::size_t q = 1; q < size(); ++q) m = m < data_[q] ? data_[q] : m; return m; }
  };
} // namespace libmpdataxx
