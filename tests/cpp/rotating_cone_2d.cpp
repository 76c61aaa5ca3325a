// The reference's 2D rotating-cone regression test
// (tests/paper_2015_GMD/3_rotating_cone_2d/rotating_cone_2d.cpp:22-162,
// statistics rotating_cone_stats.hpp:22-69) against the HIP backend's C++
// API.  Writes stats_<variant>_.txt into the directory argv[1].
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include <libmpdata++/concurr/hip.hpp>
#include <libmpdata++/solvers/mpdata.hpp>

using namespace libmpdataxx;

template <int opts_arg, int opts_iters>
void test(const std::string &outdir, const std::string &name)
{
  struct ct_params_t : ct_params_default_t
  {
    using real_t = double;
    enum { n_dims = 2 };
    enum { n_eqns = 1 };
    enum { opts = opts_arg };
  };
  const double dt = .1, dx = 1, dy = 1, omega = .1, h = 4., h0 = 1;
  int nt = 628 * 6;

  using slv_t = solvers::mpdata<ct_params_t>;
  typename slv_t::rt_params_t p;
  p.dt = dt;
  p.n_iters = opts_iters;
  p.grid_size = {101, 101};

  concurr::hip<slv_t, bcond::open, bcond::open, bcond::open, bcond::open> run(p);
  enum { x, y };
  const double r = 15. * dx, x0 = 50 * dx, y0 = 75 * dy,
               xc = .5 * (p.grid_size[x] - 1) * dx, yc = .5 * (p.grid_size[y] - 1) * dy;
  run.advectee().fill([&](int i, int j) {
    double tmp = pow(i * dx - x0, 2) + pow(j * dy - y0, 2);
    return h0 + (tmp - pow(r, 2) <= 0 ? h - sqrt(tmp / pow(r / h, 2)) : 0.);
  });
  run.advector(x).fill([&](int i, int j) { return omega * (j * dy - yc) * dt / dx; });
  run.advector(y).fill([&](int i, int j) { return -omega * (i * dx - xc) * dt / dy; });

  std::vector<double> truth(run.advectee().size());
  {
    auto a = run.advectee();
    for (std::size_t q = 0; q < truth.size(); ++q) truth[q] = a.dataFirst()[q];
  }
  std::ofstream ofs(outdir + "/stats_" + name + "_.txt");
  ofs << std::fixed << std::setprecision(8) << std::endl;
  ofs << "timestep      = 0" << std::endl;
  ofs << "min(solution) = " << run.advectee().min() << std::endl;
  ofs << "max(solution) = " << run.advectee().max() << std::endl;
  ofs << " " << std::endl;

  run.advance(nt);

  auto fc = run.advectee();
  double s2 = 0;
  for (std::size_t q = 0; q < truth.size(); ++q) s2 += pow(truth[q] - fc.dataFirst()[q], 2);
  ofs << "timestep     = " << nt << std::endl;
  ofs << "min(psi)     = " << fc.min() << std::endl;
  ofs << "max(psi)     = " << fc.max() << std::endl;
  ofs << "rms error    = " << sqrt(s2 / 101 / 101) / nt / dt << std::endl;
  std::printf("%s done\n", name.c_str());
}

int main(int argc, char **argv)
{
  const std::string outdir = argc > 1 ? argv[1] : ".";
  try
  {
    test<0, 2>(outdir, "basic");
    test<opts::fct, 2>(outdir, "fct");
    test<opts::iga | opts::fct, 2>(outdir, "iga_fct");
  }
  catch (std::exception &e)
  {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
