/** @file
 * concurr::hip -- the MI355X concurrency backend.
 *
 * Drop-in for concurr::serial / concurr::openmp (concurr/serial.hpp:15-53,
 * concurr/openmp.hpp:21-83): same template parameters
 * <solver_t, bcxl, bcxr, bcyl, bcyr, bczl, bczr>, same public interface
 * (concurr/any.hpp, concurr_common.hpp:380-453) and the same exception texts.
 * The state lives in GPU memory; advectee()/advector() return views of host
 * mirrors.  A mirror is downloaded when it is first accessed after an
 * advance() (not after every step), and uploaded before the next advance()
 * or reduction only if a writeable view of it was handed out in between;
 * advectee_global() gives a read-only view that triggers no upload.
 * advance(nt) is one library call for all nt steps (the library polls the
 * panic flag between steps); with an output decorator it is one call per
 * stretch of steps between records (with var_dt one call, the library calling
 * the decorator back for each record).
 * All work goes through the C ABI of include/mpdx.h; a solver decorated with
 * output::hdf5 / output::hdf5_xdmf records through include/mpdx_h5.h.
 */
#pragma once

#include <algorithm>
#include <array>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include <mpdx.h>

#include <libmpdata++/bcond/bcond.hpp>
#include <libmpdata++/concurr/any.hpp>
#include <libmpdata++/solvers/mpdata.hpp>

namespace libmpdataxx
{
  namespace concurr
  {
    namespace detail
    {
      // solvers decorated with an output (output/hdf5.hpp) carry a recorder type
      template <class S, class = void> struct has_output : std::false_type {};
      template <class S> struct has_output<S, std::void_t<typename S::output_recorder_t>> : std::true_type {};
      struct no_output {};
      template <class S, bool = has_output<S>::value> struct recorder_of { using type = no_output; };
      template <class S> struct recorder_of<S, true> { using type = typename S::output_recorder_t; };
    } // namespace detail

    template <class solver_t_,
              bcond::bcond_e bcxl, bcond::bcond_e bcxr,
              bcond::bcond_e bcyl = bcond::null, bcond::bcond_e bcyr = bcond::null,
              bcond::bcond_e bczl = bcond::null, bcond::bcond_e bczr = bcond::null>
    class hip : public any<typename solver_t_::real_t, solver_t_::n_dims, typename solver_t_::advance_arg_t>
    {
      public:

      using solver_t = solver_t_;
      using real_t = typename solver_t::real_t;
      using ct_params_t = typename solver_t::ct_params_t_;
      using advance_arg_t = typename solver_t::advance_arg_t;
      enum { n_dims = solver_t::n_dims, n_eqns = solver_t::n_eqns };
      using arr_t = libmpdataxx::arr_t<real_t, n_dims>;

      // concurr_common.hpp:134-148
      static_assert(n_dims == 3 || (n_dims == 2 && bczl == bcond::null && bczr == bcond::null),
                    "more boundary conditions than dimensions");

      private:

      mpdx_solver *h_ = nullptr;
      // host mirror states: stale (the device is newer), clean (equal to
      // the device), lent (a writeable view was handed out: upload before
      // the device uses the array again)
      enum mirror_state { stale, clean, lent };
      mutable std::vector<std::vector<real_t>> psi_host_;
      mutable std::vector<mirror_state> psi_state_;
      mutable std::array<std::vector<real_t>, 3> gc_host_;
      mutable std::array<mirror_state, 3> gc_state_{};
      bool panic_ = false;
      static constexpr bool with_output = detail::has_output<solver_t_>::value;
      typename detail::recorder_of<solver_t_>::type out_;

      static int bc_code(bcond::bcond_e b)
      {
        if (b == bcond::cyclic) return MPDX_BC_CYCLIC;
        if (b == bcond::open) return MPDX_BC_OPEN;
        throw std::runtime_error("libmpdata++: boundary condition not supported by the HIP backend");
      }

      void check(int rc) const
      {
        if (rc != MPDX_OK) throw std::runtime_error(mpdx_last_error(h_));
      }

      std::array<int, n_dims> interior_extent() const
      {
        std::array<int, n_dims> e{};
        for (int d = 0; d < n_dims; ++d) e[d] = grid_[d];
        return e;
      }

      static void dense_strides(const std::array<int, n_dims> &ext, int64_t st[3])
      {
        st[0] = st[1] = st[2] = 0;
        int64_t s = 1;
        for (int d = n_dims - 1; d >= 0; --d) { st[d] = s; s *= ext[d]; }
      }

      std::array<int, n_dims> grid_{};

      void advector_strides(int d, int64_t st[3]) const
      {
        int ext3[3];
        check(mpdx_advector_extents(h_, d, ext3));
        std::array<int, n_dims> ext{};
        for (int q = 0; q < n_dims; ++q) ext[q] = ext3[q];
        dense_strides(ext, st);
      }

      // mirrors handed out writeable go back to the device
      void upload() const
      {
        int64_t st[3];
        for (int e = 0; e < n_eqns; ++e)
          if (psi_state_[e] == lent)
          {
            dense_strides(interior_extent(), st);
            check(mpdx_set_advectee(h_, e, psi_host_[e].data(), st, 0));
            psi_state_[e] = clean;
          }
        for (int d = 0; d < n_dims; ++d)
          if (gc_state_[d] == lent)
          {
            advector_strides(d, st);
            check(mpdx_set_advector(h_, d, gc_host_[d].data(), st, 0));
            gc_state_[d] = clean;
          }
      }

      // the mirror of equation e / component d, current
      void fetch_advectee(int e) const
      {
        if (psi_state_.at(e) != stale) return;
        int64_t st[3];
        dense_strides(interior_extent(), st);
        check(mpdx_get_advectee(h_, e, psi_host_[e].data(), st, 0));
        psi_state_[e] = clean;
      }

      void fetch_advector(int d) const
      {
        if (gc_state_.at(d) != stale) return;
        int64_t st[3];
        advector_strides(d, st);
        check(mpdx_get_advector(h_, d, gc_host_[d].data(), st, 0));
        gc_state_[d] = clean;
      }

      // after an advance the device holds newer state (the advector too: a
      // var_dt rescale or a calc_gc hook may have changed it)
      void device_advanced()
      {
        for (auto &st : psi_state_) st = stale;
        for (int d = 0; d < n_dims; ++d) gc_state_[d] = stale;
      }

      public:

      // concurr_common ctor (concurr_common.hpp:176-187) + solver ctor checks
      explicit hip(const typename solver_t::rt_params_t &p_)
      {
        typename solver_t::rt_params_t p = p_;
        if constexpr (with_output) out_.init(p);
        mpdx_params mp;
        mpdx_params_default(&mp);
        mp.n_dims = n_dims;
        for (int d = 0; d < n_dims; ++d) { mp.grid[d] = p.grid_size[d]; grid_[d] = p.grid_size[d]; }
        mp.real_bytes = sizeof(real_t);
        mp.opts = (uint32_t)ct_params_t::opts;
        mp.n_iters = p.n_iters;
        mp.n_eqns = n_eqns;
        const bcond::bcond_e bcs[3][2] = {{bcxl, bcxr}, {bcyl, bcyr}, {bczl, bczr}};
        for (int d = 0; d < n_dims; ++d)
          for (int s = 0; s < 2; ++s) mp.bc[d][s] = bc_code(bcs[d][s]);
        mp.upwind_filter_freq = p.upwind_filter_freq;
        for (int e = 0; e < n_eqns && e < MPDX_MAX_EQNS; ++e) mp.hint_scale[e] = ct_params_t::hint_scale(e);
        mp.dt = p.dt;
        mp.max_abs_div_eps = p.max_abs_div_eps;
        mp.var_dt = ct_params_t::var_dt ? 1 : 0;
        mp.max_courant = p.max_courant;
        if (mpdx_create(&mp, &h_) != MPDX_OK) throw std::runtime_error(mpdx_last_error(nullptr));
        if constexpr (with_output)
        {
          try { out_.attach(h_); }
          catch (...) { mpdx_destroy(h_); throw; }
        }
        std::size_t n = 1;
        for (int d = 0; d < n_dims; ++d) n *= grid_[d];
        psi_host_.assign(n_eqns, std::vector<real_t>(n, real_t(0)));
        psi_state_.assign(n_eqns, stale);
        // the library polls this object's panic flag between steps
        check(mpdx_watch_panic(h_, reinterpret_cast<const volatile unsigned char *>(&panic_)));
        for (int d = 0; d < n_dims; ++d)
        {
          int ext3[3];
          check(mpdx_advector_extents(h_, d, ext3));
          std::size_t m = 1;
          for (int q = 0; q < n_dims; ++q) m *= ext3[q];
          gc_host_[d].assign(m, real_t(0));
          gc_state_[d] = stale;
        }
      }

      hip(const hip &) = delete;
      hip &operator=(const hip &) = delete;
      ~hip() { mpdx_destroy(h_); }

      // concurr_common::advance (concurr_common.hpp:380-385); the panic flag is
      // polled between time steps like solver_common::solve (solver_common.hpp:301-302)
      // with var_dt the argument is a duration (solver_common.hpp:270-274)
      void advance(advance_arg_t nt) override
      {
        upload();
        if constexpr (ct_params_t::var_dt)
        {
          // at least one step, and out_intrp_ord - 1 more once time has
          // reached the end (additional_steps, solver_common.hpp:293-294,352);
          // an output decorator records from the library's callbacks
          static_assert(ct_params_t::out_intrp_ord == 1 || ct_params_t::out_intrp_ord == 2,
                        "out_intrp_ord is 1 or 2 (output_common.hpp:78-79)");
          for (int call = 0; call < ct_params_t::out_intrp_ord && !panic_; ++call)
          {
            const int rc = call == 0 ? mpdx_advance_time(h_, nt) : mpdx_advance(h_, 1);
            if constexpr (with_output) out_.rethrow();
            check(rc);
          }
        }
        else if constexpr (with_output)
        {
          // hook_ante_loop (output_common.hpp:47-62: const.h5 and a record)
          // runs only when the solver is at timestep 0 (solver_common.hpp:277-285);
          // records between steps (hook_post_step, output_common.hpp:145-156):
          // one library call per stretch of steps up to the next record
          int64_t ts_start;
          mpdx_timestep(h_, &ts_start);
          if (ts_start == 0) out_.start(h_, nt);
          int64_t left = int64_t(nt);
          while (left > 0 && !panic_)
          {
            int64_t ts0, ts1;
            mpdx_timestep(h_, &ts0);
            const int64_t k = std::min<int64_t>(left, out_.steps_to_next_record(h_));
            check(mpdx_advance(h_, k));
            mpdx_timestep(h_, &ts1);
            left -= ts1 - ts0;
            if (ts1 - ts0 < k) break;      // stopped by the panic flag
            out_.post_step(h_);
          }
        }
        else
        {
          check(mpdx_advance(h_, int64_t(nt)));
        }
        device_advanced();
      }

      // sharedmem.hpp:522-531: interior of the current time level, base 0
      arr_t advectee(int e = 0) override
      {
        fetch_advectee(e);
        psi_state_[e] = lent;
        std::array<int, n_dims> lb{};
        return arr_t(psi_host_[e].data(), lb, interior_extent());
      }

      // read-only: no upload follows
      const arr_t advectee_global(int e = 0) override
      {
        fetch_advectee(e);
        std::array<int, n_dims> lb{};
        return arr_t(psi_host_[e].data(), lb, interior_extent());
      }

      void advectee_global_set(const arr_t arr, int e = 0) override
      {
        std::array<int, n_dims> lb{};
        arr_t(psi_host_.at(e).data(), lb, interior_extent()).assign(arr);
        psi_state_[e] = lent;
      }

      // sharedmem.hpp:543-577: faces -1..n_d-2 along d (index i <-> face i+1/2),
      // x from -1 for the components d > 0 in 3D
      arr_t advector(int d = 0) override
      {
        int ext3[3];
        check(mpdx_advector_extents(h_, d, ext3));
        std::array<int, n_dims> lb{}, ext{};
        for (int q = 0; q < n_dims; ++q) ext[q] = ext3[q];
        lb[0] = -1;
        fetch_advector(d);
        gc_state_[d] = lent;
        return arr_t(gc_host_[d].data(), lb, ext);
      }

      // sharedmem.hpp:579-591 (nug is not supported here)
      arr_t g_factor() override
      {
        throw std::runtime_error("libmpdata++: g_factor() called with nug option unset?");
      }

      bool *panic_ptr() override { return &panic_; }

      const real_t time() const override
      {
        double t;
        mpdx_time(h_, &t);
        return real_t(t);
      }

      const real_t min(int e = 0) const override
      {
        double v;
        upload();
        check(mpdx_reduce(h_, e, MPDX_RED_MIN, &v));
        return real_t(v);
      }

      const real_t max(int e = 0) const override
      {
        double v;
        upload();
        check(mpdx_reduce(h_, e, MPDX_RED_MAX, &v));
        return real_t(v);
      }

      // FP64 Kahan-compensated sum of the advectee (sharedmem.hpp:115-155 with sum_khn)
      double mass(int e = 0) const
      {
        double v;
        upload();
        check(mpdx_reduce(h_, e, MPDX_RED_SUM, &v));
        return v;
      }

      mpdx_solver *handle() const { return h_; }
    };
  } // namespace concurr
} // namespace libmpdataxx
