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
 * at least outstart (hook_post_step, output_common.hpp:145-156).  Output with
 * var_dt (interpolated records, output_common.hpp:72-140) is not implemented.
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

        mpdx_h5_setup s_{};
        std::map<int, info_t> outvars_;
        std::string outdir_;
        advance_arg_t outfreq_ = 1, outstart_ = 0;
        int outwindow_ = 1;
        std::vector<int64_t> records_;   // timesteps already in temp.xmf
        std::vector<std::vector<real_t>> buf_;

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
          static_assert(!solver_t::ct_params_t_::var_dt, "output with var_dt is not implemented by the HIP backend");
          outfreq_ = p.outfreq;
          outstart_ = p.outstart;
          outwindow_ = p.outwindow;
          outvars_ = p.outvars;
          outdir_ = p.outdir;
          if (int(outfreq_) <= 0 || int(outstart_) % int(outfreq_) != 0)
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
          s_.var_dt = 0;
          s_.max_courant = p.max_courant;
          s_.outfreq = double(outfreq_);
          s_.outstart = double(outstart_);
          s_.outwindow = outwindow_;
          s_.xdmf = xdmf ? 1 : 0;
          std::size_t n = 1;
          for (int d = 0; d < n_dims; ++d) n *= std::size_t(p.grid_size[d]);
          buf_.assign(outvars_.size(), std::vector<real_t>(n));
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

        void record(mpdx_solver *h)
        {
          int64_t ts;
          mpdx_timestep(h, &ts);
          // time = timestep * dt in the solver's real_t (solver_common.hpp:348)
          const real_t record_time = real_t(ts) * real_t(s_.dt);
          int64_t st[3] = {0, 0, 0};
          int64_t s = 1;
          for (int d = n_dims - 1; d >= 0; --d) { st[d] = s; s *= s_.grid[d]; }
          std::vector<const char *> names;
          std::vector<const void *> fields;
          std::size_t q = 0;
          for (const auto &v : outvars_)
          {
            if (mpdx_get_advectee(h, v.first, buf_[q].data(), st, 0) != MPDX_OK)
              throw std::runtime_error(mpdx_last_error(h));
            names.push_back(v.second.name.c_str());
            fields.push_back(buf_[q].data());
            ++q;
          }
          check(mpdx_h5_record(&s_, outdir_.c_str(), ts, double(record_time), int(names.size()), names.data(),
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
