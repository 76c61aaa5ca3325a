// HDF5/XDMF output of a var_dt solver through the C++ API: records at
// multiples of outfreq interpolated in time (output_common.hpp:72-140) with
// out_intrp_ord 1 and 2.  2D, 40 x 30, open bounds, a cone in solid-body
// rotation, Courant number rescaled to max_courant = 0.4
// (solver_common.hpp:199-210), outfreq = 0.5, two advance calls of 1.3 and
// 0.9.  For each ord it writes into <argv[1]>/ord<ord>/ the decorator's files
// (const.h5, timestepNNNNNNNNNN.{h5,xmf}, temp.xmf) and in.f64: the advectee
// and the two advector components before the first call (row-major doubles).
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include <libmpdata++/concurr/hip.hpp>
#include <libmpdata++/output/hdf5_xdmf.hpp>
#include <libmpdata++/solvers/mpdata.hpp>

using namespace libmpdataxx;

template <int ord>
void run_case(const std::string &dir)
{
  struct ct_params_t : ct_params_default_t
  {
    using real_t = double;
    enum { n_dims = 2 };
    enum { n_eqns = 1 };
    enum { opts = opts::fct };
    enum { var_dt = true };
    enum { out_intrp_ord = ord };
  };
  using slv_t = solvers::mpdata<ct_params_t>;
  using slv_out_t = output::hdf5_xdmf<slv_t>;
  typename slv_out_t::rt_params_t p;
  p.dt = 0.3;
  p.n_iters = 2;
  p.grid_size = {40, 30};
  p.max_courant = 0.4;
  p.outfreq = 0.5;
  p.outvars[0].name = "psi";
  p.outdir = dir;

  concurr::hip<slv_out_t, bcond::open, bcond::open, bcond::open, bcond::open> run(p);
  const double xc = 19.5, yc = 14.5, x0 = 12., y0 = 15., r = 6., omega = 0.05;
  run.advectee().fill([&](int i, int j) {
    const double d2 = (i - x0) * (i - x0) + (j - y0) * (j - y0);
    return 1. + (d2 <= r * r ? 4. - 4. * std::sqrt(d2) / r : 0.);
  });
  run.advector(0).fill([&](int i, int j) { return omega * (j - yc) * p.dt; });
  run.advector(1).fill([&](int i, int j) { return -omega * (i - xc) * p.dt; });

  // the directory exists once the first advance has written const.h5; the
  // initial state goes to a buffer until then
  std::vector<double> in;
  {
    auto a = run.advectee();
    in.insert(in.end(), a.dataFirst(), a.dataFirst() + a.size());
    for (int d = 0; d < 2; ++d)
    {
      auto c = run.advector(d);
      in.insert(in.end(), c.dataFirst(), c.dataFirst() + c.size());
    }
  }
  run.advance(1.3);
  run.advance(0.9);
  std::FILE *g = std::fopen((dir + "/in.f64").c_str(), "wb");
  std::fwrite(in.data(), sizeof(double), in.size(), g);
  std::fclose(g);
  std::printf("%s written, time = %.6f\n", dir.c_str(), double(run.time()));
}

int main(int argc, char **argv)
{
  const std::string outdir = argc > 1 ? argv[1] : ".";
  try
  {
    run_case<1>(outdir + "/ord1");
    run_case<2>(outdir + "/ord2");
  }
  catch (const std::exception &e)
  {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
