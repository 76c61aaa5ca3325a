// A var_dt solver driven through the C++ API: advance(duration) with
// out_intrp_ord 1 and 2 (solver_common.hpp:292-355: steps while time < end
// or additional_steps > 0, additional_steps starting at out_intrp_ord and
// dropping on each step that ends at or past the end).  2D, 40 x 30, open
// bounds, a cone in solid-body rotation with a Courant number rescaled to
// max_courant = 0.4 (solver_common.hpp:199-210).  For each ord it calls
// advance(0), advance(1.3), advance(0.25) and writes, into the directory
// argv[1], var_dt_ord<ord>_in.f64 (the advectee and the two advector
// components before the first call, row-major doubles), var_dt_ord<ord>.txt
// (the time after each call, %.17g) and var_dt_ord<ord>.f64 (the final
// field).
#include <cmath>
#include <cstdio>
#include <string>

#include <libmpdata++/concurr/hip.hpp>
#include <libmpdata++/solvers/mpdata.hpp>

using namespace libmpdataxx;

template <int ord>
void run_case(const std::string &outdir)
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
  typename slv_t::rt_params_t p;
  p.dt = 0.3;
  p.n_iters = 2;
  p.grid_size = {40, 30};
  p.max_courant = 0.4;

  concurr::hip<slv_t, bcond::open, bcond::open, bcond::open, bcond::open> run(p);
  const double xc = 19.5, yc = 14.5, x0 = 12., y0 = 15., r = 6., omega = 0.05;
  run.advectee().fill([&](int i, int j) {
    const double d2 = (i - x0) * (i - x0) + (j - y0) * (j - y0);
    return 1. + (d2 <= r * r ? 4. - 4. * std::sqrt(d2) / r : 0.);
  });
  run.advector(0).fill([&](int i, int j) { return omega * (j - yc) * p.dt; });
  run.advector(1).fill([&](int i, int j) { return -omega * (i - xc) * p.dt; });

  const std::string base = outdir + "/var_dt_ord" + std::to_string(ord);
  // the initial state as the C++ API holds it: the advectee, then each
  // advector component over its whole extent, row-major
  {
    std::FILE *g = std::fopen((base + "_in.f64").c_str(), "wb");
    auto a = run.advectee();
    std::fwrite(a.dataFirst(), sizeof(double), a.size(), g);
    for (int d = 0; d < 2; ++d)
    {
      auto c = run.advector(d);
      std::fwrite(c.dataFirst(), sizeof(double), c.size(), g);
    }
    std::fclose(g);
  }
  std::FILE *f = std::fopen((base + ".txt").c_str(), "w");
  for (double dur : {0., 1.3, .25})
  {
    run.advance(dur);
    std::fprintf(f, "%.17g\n", double(run.time()));
  }
  std::fclose(f);
  auto a = run.advectee();
  std::FILE *g = std::fopen((base + ".f64").c_str(), "wb");
  std::fwrite(a.dataFirst(), sizeof(double), a.size(), g);
  std::fclose(g);
}

int main(int argc, char **argv)
{
  const std::string outdir = argc > 1 ? argv[1] : ".";
  run_case<1>(outdir);
  run_case<2>(outdir);
  return 0;
}
