/** @file
 * output::hdf5<solver_t> and output::hdf5_xdmf<solver_t> for the HIP backend.
 *
 * Same decorator names, rt_params_t fields (outfreq, outstart, outwindow,
 * outvars, outdir; output/detail/output_common.hpp:169-177) and file layout
 * as the reference's (output/hdf5.hpp, output/hdf5_xdmf.hpp), so that
 *
 *   concurr::hip<output::hdf5_xdmf<solvers::mpdata<ct_params_t>>, ...>
 *
 * writes const.h5, timestepNNNNNNNNNN.h5/.xmf and temp.xmf the way
 * concurr::serial<output::hdf5_xdmf<...>> does.  The decorator only
 * describes the output; concurr::hip records through the C ABI of
 * include/mpdx_h5.h (link libmpdx_h5 and HDF5).
 *
 * Records are taken when the reference takes them: at the start of the first
 * advance(), from timestep 0 (hook_ante_loop, output_common.hpp:47-62, which
 * solver_common.hpp:277-285 runs only then; it also writes const.h5 with that
 * call's nt), and after each step
 * whose number, less some t in [0, outwindow), is a multiple of outfreq and
 * at least outstart (hook_post_step, output_common.hpp:145-156).  With var_dt
 * the library keeps the reference's interpolation schedule
 * (output_common.hpp:72-140, mpdx_set_var_dt_output) and calls the recorder
 * back: at the start (after dt was rescaled, so const.h5 carries that dt) and
 * for each record at k * outfreq, time-interpolated to out_intrp_ord 1 or 2.
 */
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include <mpdx.h>
#include <mpdx_h5.h>

namespace libmpdataxx
{
  namespace output
  {
    namespace detail
    {
      struct info_t { std::string name, unit; };

      // the recorder concurr::hip keeps for a decorated solver
      template <class solver_t, bool xdmf>
      class recorder
      {
        using real_t = typename solver_t::real_t;
        using advance_arg_t = typename solver_t::advance_arg_t;
        enum { n_dims = solver_t::n_dims, n_eqns = solver_t::n_eqns };
        static constexpr bool var_dt = solver_t::ct_params_t_::var_dt;

        mpdx_h5_setup s_{};
        std::map<int, info_t> outvars_;
        std::string outdir_;
        advance_arg_t outfreq_ = 1, outstart_ = 0;
        int outwindow_ = 1;
        std::vector<int64_t> records_;   // timesteps already in temp.xmf
        std::vector<std::vector<real_t>> buf_;
        std::string cb_err_;             // exception text from a library callback

        static void check(int rc)
        {
          if (rc != 0) throw std::runtime_error(std::string("libmpdata++: ") + mpdx_h5_last_error());
        }

        public:

        // output_common ctor (output_common.hpp:181-208); dt/di/dj/dk of 0
        // become 1 "for output purposes", which the caller applies to the solver too
        template <class rt_params_t>
        void init(rt_params_t &p)
        {
          outfreq_ = p.outfreq;
          outstart_ = p.outstart;
          outwindow_ = p.outwindow;
          outvars_ = p.outvars;
          outdir_ = p.outdir;
          if constexpr (var_dt)
          {
            // output_common.hpp:193-194
            if (!(outfreq_ > 0)) throw std::runtime_error("libmpdata++: outfreq must be positive");
            if (outstart_ != 0) throw std::runtime_error("libmpdata++: outstart must be 0 with var_dt");
          }
          else if (int(outfreq_) <= 0 || int(outstart_) % int(outfreq_) != 0)
            throw std::runtime_error("libmpdata++: outstart must be a multiple of outfreq");
          if (outvars_.empty() && n_eqns == 1) outvars_ = {{0, {"", ""}}};
          for (const auto &v : outvars_)
            if (v.first < 0 || v.first >= n_eqns) throw std::runtime_error("libmpdata++: outvars index out of range");
          if (p.dt == 0) p.dt = 1;
          if (p.di == 0) p.di = 1;
          if (p.dj == 0) p.dj = 1;
          if (p.dk == 0) p.dk = 1;
          s_.n_dims = n_dims;
          for (int d = 0; d < n_dims; ++d) s_.grid[d] = p.grid_size[d];
          s_.real_bytes = sizeof(real_t);
          s_.dijk[0] = p.di; s_.dijk[1] = p.dj; s_.dijk[2] = p.dk;
          s_.dt = p.dt;
          s_.n_iters = p.n_iters;
          s_.opts = (uint32_t)solver_t::ct_params_t_::opts;
          s_.var_dt = var_dt ? 1 : 0;
          s_.max_courant = p.max_courant;
          s_.outfreq = double(outfreq_);
          s_.outstart = double(outstart_);
          s_.outwindow = outwindow_;
          s_.xdmf = xdmf ? 1 : 0;
          std::size_t n = 1;
          for (int d = 0; d < n_dims; ++d) n *= std::size_t(p.grid_size[d]);
          buf_.assign(outvars_.size(), std::vector<real_t>(n));
        }

        // var_dt: the library runs the interpolation schedule and calls back
        void attach(mpdx_solver *h)
        {
          if constexpr (var_dt)
          {
            std::vector<int> vars;
            for (const auto &v : outvars_) vars.push_back(v.first);
            if (mpdx_set_var_dt_output(h, double(outfreq_), solver_t::ct_params_t_::out_intrp_ord, int(vars.size()),
                                       vars.data(), &recorder::on_output, this) != MPDX_OK)
              throw std::runtime_error(mpdx_last_error(h));
          }
        }

        // an exception a callback turned into an error code
        void rethrow()
        {
          if (cb_err_.empty()) return;
          const std::string m = cb_err_;
          cb_err_.clear();
          throw std::runtime_error(m);
        }

        // hdf5.hpp:66-235 start() followed by record_all()
        void start(mpdx_solver *h, advance_arg_t nt)
        {
          check(mpdx_h5_start(&s_, outdir_.c_str(), double(nt)));
          record(h);
        }

        // hook_post_step (output_common.hpp:150-155): one record per matching t
        void post_step(mpdx_solver *h)
        {
          int64_t ts;
          mpdx_timestep(h, &ts);
          for (int t = 0; t < outwindow_; ++t)
            if ((ts - t) % int64_t(outfreq_) == 0 && (ts - t) >= int64_t(outstart_)) record(h);
        }

        // steps from the current one to the next step whose post_step
        // records (at least 1), so that the steps in between can go to the
        // library in one call
        int64_t steps_to_next_record(mpdx_solver *h) const
        {
          int64_t ts;
          mpdx_timestep(h, &ts);
          const int64_t f = int64_t(outfreq_), horizon = f + int64_t(outstart_) + outwindow_ + 1;
          for (int64_t k = 1; k < horizon; ++k)
            for (int t = 0; t < outwindow_; ++t)
              if ((ts + k - t) % f == 0 && (ts + k - t) >= int64_t(outstart_)) return k;
          return horizon;
        }

        // with var_dt, the start (hook_ante_loop, output_common.hpp:47-62:
        // const.h5 with the rescaled dt, then a record at the current time)
        // and each interpolated record (output_common.hpp:104-156)
        static int on_output(mpdx_solver *h, int kind, double t, void *user)
        {
          auto *r = static_cast<recorder *>(user);
          try
          {
            double rt = t;
            if (kind == MPDX_OUT_START)
            {
              mpdx_dt(h, &r->s_.dt);
              r->check(mpdx_h5_start(&r->s_, r->outdir_.c_str(), double(real_t(t))));
              mpdx_time(h, &rt);
            }
            r->record(h, true, double(real_t(rt)));
            return 0;
          }
          catch (const std::exception &e)
          {
            r->cb_err_ = e.what();
            return 1;
          }
        }

        // interp: the library's interpolation buffers at time t (var_dt),
        // else the current state at timestep * dt
        void record(mpdx_solver *h, bool interp = false, double t = 0)
        {
          int64_t ts;
          mpdx_timestep(h, &ts);
          // time = timestep * dt in the solver's real_t (solver_common.hpp:348)
          const double record_time = interp ? t : double(real_t(ts) * real_t(s_.dt));
          int64_t st[3] = {0, 0, 0};
          int64_t s = 1;
          for (int d = n_dims - 1; d >= 0; --d) { st[d] = s; s *= s_.grid[d]; }
          std::vector<const char *> names;
          std::vector<const void *> fields;
          std::size_t q = 0;
          for (const auto &v : outvars_)
          {
            if ((interp ? mpdx_get_output(h, v.first, buf_[q].data(), st, 0)
                        : mpdx_get_advectee(h, v.first, buf_[q].data(), st, 0)) != MPDX_OK)
              throw std::runtime_error(mpdx_last_error(h));
            names.push_back(v.second.name.c_str());
            fields.push_back(buf_[q].data());
            ++q;
          }
          check(mpdx_h5_record(&s_, outdir_.c_str(), ts, record_time, int(names.size()), names.data(),
                               fields.data(), int(records_.size()), records_.data()));
          if (xdmf) records_.push_back(ts);
        }
      };

      template <class solver_t>
      struct output_params : solver_t::rt_params_t
      {
        typename solver_t::advance_arg_t outfreq = 1, outstart = 0;
        int outwindow = 1;
        std::map<int, info_t> outvars;
        std::string outdir;
      };
    } // namespace detail

    // output/hdf5.hpp: HDF5 files only
    template <class solver_t>
    class hdf5 : public solver_t
    {
      public:
      using info_t = detail::info_t;
      using rt_params_t = detail::output_params<solver_t>;
      using output_recorder_t = detail::recorder<solver_t, false>;
    };

    // output/hdf5_xdmf.hpp: HDF5 files plus the XDMF markup (2D and 3D)
    template <class solver_t>
    class hdf5_xdmf : public solver_t
    {
      public:
      static_assert(solver_t::n_dims > 1, "only 2D and 3D output supported");
      using info_t = detail::info_t;
      using rt_params_t = detail::output_params<solver_t>;
      using output_recorder_t = detail::recorder<solver_t, true>;
    };
  } // namespace output
} // namespace libmpdataxx
