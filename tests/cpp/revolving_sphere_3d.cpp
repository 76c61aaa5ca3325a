// The reference's 3D revolving-sphere regression test
// (tests/paper_2015_GMD/4_revolving_sphere_3d/revolving_sphere_3d.cpp:17-101,
// statistics revolving_sphere_stats.hpp:22-65) written against the HIP
// backend's C++ API: only the concurrency backend (concurr::hip) and the
// array-filling syntax differ.  Writes stats_<variant>.txt and the final
// field as raw float32 <variant>_556.f32 into the directory argv[1].
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
  enum { x, y, z };
  struct ct_params_t : ct_params_default_t
  {
    using real_t = double;
    enum { n_dims = 3 };
    enum { n_eqns = 1 };
    enum { opts = opts_arg };
  };

  int nt = 556;
  int nx = 59;

  using slv_t = solvers::mpdata<ct_params_t>;
  typename slv_t::rt_params_t p;
  p.n_iters = opts_iters;
  p.grid_size = {nx, nx, nx};

  const double
    dt = 0.018 * 2 * M_PI,
    L = 100,
    dx = L / (nx - 1),
    dy = dx,
    dz = dx,
    h = 4,
    r = 15,
    d = 25 / sqrt(3),
    x0 = 50 - d,
    y0 = 50 + d,
    z0 = 50 + d;
  p.di = dx; p.dj = dy; p.dk = dz; p.dt = dt;

  concurr::hip<slv_t, bcond::open, bcond::open, bcond::open, bcond::open, bcond::open, bcond::open> slv(p);

  slv.advectee().fill([&](int i, int j, int k) {
    double tmp = pow(i * dx - x0, 2) + pow(j * dx - y0, 2) + pow(k * dx - z0, 2);
    return tmp - pow(r, 2) <= 0 ? h : 0.;
  });

  const double omega = 0.1, xc = 50, yc = 50, zc = 50;
  slv.advector(x).fill([&](int i, int j, int k) { return omega / sqrt(3) * (-(j * dy - yc) + (k * dz - zc)) * dt / dx; });
  slv.advector(y).fill([&](int i, int j, int k) { return omega / sqrt(3) * ( (i * dx - xc) - (k * dz - zc)) * dt / dy; });
  slv.advector(z).fill([&](int i, int j, int k) { return omega / sqrt(3) * (-(i * dx - xc) + (j * dy - yc)) * dt / dz; });

  std::vector<double> ic(slv.advectee().size());
  {
    auto a = slv.advectee();
    for (std::size_t q = 0; q < ic.size(); ++q) ic[q] = a.dataFirst()[q];
  }
  std::ofstream ofs(outdir + "/stats_" + name + ".txt");
  ofs << std::fixed << std::setprecision(8) << std::endl;
  ofs << "timestep      = 0" << std::endl;
  ofs << "min(solution) = " << slv.advectee().min() << std::endl;
  ofs << "max(solution) = " << slv.advectee().max() << std::endl;
  ofs << " " << std::endl;

  slv.advance(nt);

  auto fc = slv.advectee();
  double linf = 0, s2 = 0, s1 = 0;
  for (std::size_t q = 0; q < ic.size(); ++q)
  {
    const double df = fc.dataFirst()[q] - ic[q];
    linf = std::max(linf, std::abs(df));
    s2 += df * df;
    s1 += std::abs(df);
  }
  const double n3 = pow(nx, 3);
  ofs << "timestep = " << nt << std::endl;
  ofs << "max(solution) = " << fc.max() << std::endl;
  ofs << "min(solution) = " << fc.min() << std::endl;
  ofs << "Linf = " << linf << std::endl;
  ofs << "L2 = " << 1.0 / (nt * dt) * sqrt(1.0 / n3 * s2) << std::endl;
  ofs << "L1 = " << 1.0 / (nt * dt) * (1.0 / n3 * s1) << std::endl;

  std::vector<float> f32(fc.size());
  for (std::size_t q = 0; q < f32.size(); ++q) f32[q] = float(fc.dataFirst()[q]);
  std::ofstream bin(outdir + "/" + name + "_556.f32", std::ios::binary);
  bin.write(reinterpret_cast<const char *>(f32.data()), f32.size() * sizeof(float));
  std::printf("%s done, time = %.6f\n", name.c_str(), double(slv.time()));
}

int main(int argc, char **argv)
{
  const std::string outdir = argc > 1 ? argv[1] : ".";
  try
  {
    test<0, 1>(outdir, "upwind");
    test<opts::iga, 2>(outdir, "iga");
    test<0, 2>(outdir, "basic");
    test<opts::iga | opts::fct, 2>(outdir, "iga_fct");
    test<opts::fct, 2>(outdir, "fct");
  }
  catch (std::exception &e)
  {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
