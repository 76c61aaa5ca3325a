// The reference's 3D revolving-sphere run with its HDF5/XDMF output
// (tests/paper_2015_GMD/4_revolving_sphere_3d/revolving_sphere_3d.cpp:17-101:
// output::hdf5_xdmf over solvers::mpdata, outfreq = nt, outvars {0: psi})
// written against the HIP backend's C++ API.  For each variant it writes the
// reference's refdata layout into <argv[1]>/<variant>/: const.h5,
// timestep0000000000.{h5,xmf}, timestep0000000556.{h5,xmf}, temp.xmf.
#include <cmath>
#include <cstdio>
#include <string>

#include <libmpdata++/concurr/hip.hpp>
#include <libmpdata++/output/hdf5_xdmf.hpp>
#include <libmpdata++/solvers/mpdata.hpp>

using namespace libmpdataxx;

template <int opts_arg, int opts_iters>
void test(const std::string &dir_name)
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
  using slv_out_t = output::hdf5_xdmf<slv_t>;
  typename slv_out_t::rt_params_t p;

  p.n_iters = opts_iters;
  p.grid_size = {nx, nx, nx};
  p.outfreq = nt;
  p.outvars[0].name = "psi";
  p.outdir = dir_name;

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

  concurr::hip<slv_out_t, bcond::open, bcond::open, bcond::open, bcond::open, bcond::open, bcond::open> slv(p);

  slv.advectee().fill([&](int i, int j, int k) {
    double tmp = pow(i * dx - x0, 2) + pow(j * dx - y0, 2) + pow(k * dx - z0, 2);
    return tmp - pow(r, 2) <= 0 ? h : 0.;
  });

  const double omega = 0.1, xc = 50, yc = 50, zc = 50;
  slv.advector(x).fill([&](int i, int j, int k) { return omega / sqrt(3) * (-(j * dy - yc) + (k * dz - zc)) * dt / dx; });
  slv.advector(y).fill([&](int i, int j, int k) { return omega / sqrt(3) * ( (i * dx - xc) - (k * dz - zc)) * dt / dy; });
  slv.advector(z).fill([&](int i, int j, int k) { return omega / sqrt(3) * (-(i * dx - xc) + (j * dy - yc)) * dt / dz; });

  slv.advance(nt);
  std::printf("%s written, time = %.6f\n", dir_name.c_str(), double(slv.time()));
}

int main(int argc, char **argv)
{
  const std::string outdir = argc > 1 ? argv[1] : ".";
  try
  {
    test<0, 1>(outdir + "/upwind");
    test<opts::iga, 2>(outdir + "/iga");
    test<0, 2>(outdir + "/basic");
    test<opts::iga | opts::fct, 2>(outdir + "/iga_fct");
    test<opts::fct, 2>(outdir + "/fct");
  }
  catch (std::exception &e)
  {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
