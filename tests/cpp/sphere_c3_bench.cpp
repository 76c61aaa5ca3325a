// The headline workload through the C++ drop-in API: 3D MPDATA-FCT,
// n_iters = 3, open boundaries, FP32 (real_t = float), solid-body rotation
// about z moving a discontinuous sphere, written the way a libmpdata++ user
// writes a concurr::serial program (cf. tests/paper_2015_GMD/
// 4_revolving_sphere_3d/revolving_sphere_3d.cpp:69-97) with concurr::hip in
// its place.  argv: grid size n (default 512), timed steps K (default 20),
// warm-up steps W (default 3).  Prints one JSON line: the best of two timed
// advance(K) calls (std::chrono around advance, inputs already on the GPU),
// and the FP64 mass, min and max after W + 2 K steps.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include <libmpdata++/concurr/hip.hpp>
#include <libmpdata++/solvers/mpdata.hpp>

using namespace libmpdataxx;

int main(int argc, char **argv)
{
  const int n = argc > 1 ? std::atoi(argv[1]) : 512;
  const int K = argc > 2 ? std::atoi(argv[2]) : 20;
  const int W = argc > 3 ? std::atoi(argv[3]) : 3;

  struct ct_params_t : ct_params_default_t
  {
    using real_t = float;
    enum { n_dims = 3 };
    enum { n_eqns = 1 };
    enum { opts = opts::fct };
  };
  using slv_t = solvers::mpdata<ct_params_t>;
  typename slv_t::rt_params_t p;
  p.n_iters = 3;
  p.grid_size = {n, n, n};
  concurr::hip<slv_t, bcond::open, bcond::open, bcond::open, bcond::open, bcond::open, bcond::open> run(p);

  // workloads.rotation_sphere_3d: radius 40 n / 256, centre offset 48 n / 256
  // along x, rotation rate w = 0.5 / c about the centre c = (n - 1) / 2
  const double c = (n - 1) / 2.0, w = 0.5 / c, radius = 40.0 * n / 256, offset = 48.0 * n / 256;
  run.advectee().fill([&](int i, int j, int k) {
    const double r2 = (i - (c + offset)) * (i - (c + offset)) + (j - c) * (j - c) + (k - c) * (k - c);
    return r2 <= radius * radius ? 1.0f : 0.0f;
  });
  run.advector(0).fill([&](int, int j, int) { return float(-w * (j - c)); });
  run.advector(1).fill([&](int i, int, int) { return float(w * (i - c)); });
  run.advector(2).fill([&](int, int, int) { return 0.0f; });
  const double mass0 = run.mass();

  run.advance(W);   // uploads the inputs, first-step work, graph capture
  double best = 1e300;
  for (int rep = 0; rep < 2; ++rep)
  {
    const auto t0 = std::chrono::steady_clock::now();
    run.advance(K);
    const auto t1 = std::chrono::steady_clock::now();
    best = std::min(best, std::chrono::duration<double>(t1 - t0).count());
  }
  const double cells = double(n) * n * n;
  std::printf("{\"grid\": %d, \"steps\": %d, \"warmup\": %d, \"gcell_per_s\": %.6f, \"ms_per_step\": %.6f, "
              "\"mass0\": %.17g, \"mass\": %.17g, \"min\": %.9g, \"max\": %.9g}\n",
              n, K, W, cells * K / best / 1e9, best * 1e3 / K, mass0, run.mass(), double(run.min()),
              double(run.max()));
  return 0;
}
