/** @file
 * Abstract solver interface, as concurr::any (concurr/any.hpp:15-74).
 */
#pragma once

#include <cassert>
#include <stdexcept>

#include <libmpdata++/array.hpp>

namespace libmpdataxx
{
  namespace concurr
  {
    template <typename real_t, int n_dims, typename advance_arg_t = int>
    struct any
    {
      using arr_t = libmpdataxx::arr_t<real_t, n_dims>;
      virtual void advance(advance_arg_t) = 0;
      virtual arr_t advectee(int eqn = 0) = 0;
      virtual const arr_t advectee_global(int eqn = 0) = 0;
      virtual void advectee_global_set(const arr_t, int eqn = 0) = 0;
      virtual arr_t advector(int dim = 0) = 0;
      virtual arr_t g_factor() = 0;
      virtual bool *panic_ptr() = 0;
      virtual const real_t time() const = 0;
      virtual const real_t min(int eqn = 0) const = 0;
      virtual const real_t max(int eqn = 0) const = 0;
      virtual ~any() {}
    };
  } // namespace concurr
} // namespace libmpdataxx
