/** @file
 * solvers::mpdata<ct_params_t> for the HIP backend.
 *
 * Same template signature, member types and rt_params_t fields as the
 * reference (solvers/mpdata.hpp:24-54, mpdata_common.hpp:74-78,
 * solver_common.hpp:217-221, solver_3d.hpp:438-441), so user code that
 * configures a reference solver configures this one.  The class only
 * describes the solver; concurr::hip runs it on the GPU.
 */
#pragma once

#include <array>
#include <limits>
#include <type_traits>

#include <libmpdata++/array.hpp>
#include <libmpdata++/opts.hpp>

namespace libmpdataxx
{
  namespace solvers
  {
    struct mpdata_family_tag {};

    template <typename ct_params_t, int minhalo = 0>
    class mpdata
    {
      public:

      using ct_params_t_ = ct_params_t;
      using real_t = typename ct_params_t::real_t;
      enum { n_dims = ct_params_t::n_dims };
      enum { n_eqns = ct_params_t::n_eqns };
      enum { n_tlev = 2 };
      static constexpr opts::opts_t opts = ct_params_t::opts;
      // halo width: mpdata halo 1, FCT needs 2 (mpdata_fct_common.hpp:15)
      enum { halo = opts::isset(ct_params_t::opts, opts::fct) ? 2 : 1 };
      using arr_t = libmpdataxx::arr_t<real_t, n_dims>;
      using advance_arg_t = typename std::conditional<ct_params_t::var_dt, real_t, int>::type;
      using solver_family = mpdata_family_tag;

      static_assert(n_dims == 2 || n_dims == 3, "the HIP backend implements 2D and 3D");
      static_assert(std::is_same<real_t, float>::value || std::is_same<real_t, double>::value,
                    "real_t must be float or double");
      static_assert(!(opts::isset(opts, opts::abs) && opts::isset(opts, opts::iga)),
                    "abs & iga options are mutually exclusive");

      struct rt_params_t
      {
        std::array<int, n_dims> grid_size;
        real_t dt = 0;
        real_t max_abs_div_eps = std::numeric_limits<real_t>::epsilon();
        real_t max_courant = real_t(0.5);
        real_t di = 0, dj = 0, dk = 0;
        int n_iters = 2;
        int upwind_filter_freq = 0;
      };
    };
  } // namespace solvers
} // namespace libmpdataxx
