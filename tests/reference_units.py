"""Inputs of two of the reference's unit tests that bound the FCT path,
built as the reference programs build them (real_t = float), in the views
of this backend (advectee [x][y]; advector(0) over faces -1/2 .. n_x-3/2,
advector(1) over x = -1 .. n_x-1 and faces 1/2 .. n_y-3/2).

* cone_bugs (tests/unit/cone_bugs/test_issue.cpp): 15 x 14 cyclic, FCT,
  n_iters = 2, a cone of height 4 on a background of 1 under a uniform
  Courant number 0.3 along both axes; after 2 steps the minimum is exactly 1.
* hint_scale (tests/unit/hint_scale/hint_scale.cpp): 4 x 4, cyclic in x and
  open in y, FCT | iga, n_iters = 2, psi ~ 8e7 next to zeros; with
  hint_scale = 24 one step keeps max <= the initial max and min >= 0.  The
  program's comma lists are written over the arrays' raw memory (interior
  values, then halo NaNs); the values below are the interior part of each
  row.  Its other half (values below zero without scaling) needs the
  program's fesetround(FE_DOWNWARD) float arithmetic and is not reproduced.
* convergence_2d / _3d (tests/sandbox/convergence_2d_3d/convergence_2d.cpp:
  18-92, convergence_3d.cpp:19-97): psi = 2 + product of sin(2 pi i dx) over
  the axes on an np^d cyclic grid, advector (1, 2(, 3)) dt / dx with
  dt = 0.2 dx (2D) or 0.1 dx (3D), one period; round(log2(L2 error at the
  coarse / fine np)) is 2 for basic and for iga|fct with n_iters = 2 (2D:
  33 / 65, 3D: 17 / 33).
"""
import numpy as np

F32_EPS = float(np.finfo(np.float32).eps)


def cone_bugs_case():
    nx, ny = 15, 14
    f = np.float32
    dx = dy = f(1)
    h, h0, r = f(4), f(1), f(12) * dx
    x0, y0 = f(1), f(0)
    i = np.arange(nx, dtype=f)[:, None]
    j = np.arange(ny, dtype=f)[None, :]
    d2 = (i * dx - x0) ** 2 + (j * dy - y0) ** 2
    psi = h0 + np.where(d2 - r * r <= 0, h * (f(1) - d2 / (r * r)) ** 2, f(0))
    gc = [np.full((nx, ny), f(0.3)), np.full((nx + 1, ny - 1), f(0.3))]
    return dict(grid=(nx, ny), bcs=[(0, 0), (0, 0)], n_iters=2, nt=2,
                psi=psi.astype(np.float64), gc=[g.astype(np.float64) for g in gc])


def hint_scale_case():
    f = np.float32
    psi = np.array([[0, 0, 8.16638e+07, 8.16339e+07],
                    [0, 0, 7.8215e+07, 8.15976e+07],
                    [0, 0, 8.16724e+07, 8.16339e+07],
                    [0, 0, 8.16638e+07, 8.16339e+07]], dtype=f)
    nan = np.nan
    gx = np.array([[nan, nan, nan, nan],
                   [0.000572957855183631, 0.000286478927591815, -0.000286478927591815, -0.000572957855183631],
                   [-0.00114591582678258, -0.000572957913391292, 0.000572957855183631, 0.00114591582678258],
                   [0.00057295779697597, 0.000286478898487985, -0.000286478869384154, -0.00057295779697597]],
                  dtype=f)
    gy = np.array([[nan, nan, nan],
                   [0, 0, 0],
                   [0.000859436870086938, 0.00171887374017387, 0.000859436928294599],
                   [-0.000859436753671616, -0.00171887350734323, -0.000859436870086938],
                   [-1.70754757555791e-10, -3.41509515111582e-10, -1.70754757555791e-10]], dtype=f)
    return dict(grid=(4, 4), bcs=[(0, 0), (1, 1)], n_iters=2, nt=1,
                psi=psi.astype(np.float64), gc=[gx.astype(np.float64), gy.astype(np.float64)])


def convergence_case(np_, nd=2):
    time, dx = 1.0, 1.0 / (np_ - 1)
    dy = dz = dx
    dt = (0.2 if nd == 2 else 0.1) * dx
    nt = int(time / dt)
    ax = np.arange(np_)
    if nd == 2:
        psi = 2 + np.sin(2 * np.pi * ax[:, None] * dx) * np.sin(2 * np.pi * ax[None, :] * dy)
        gc = [np.full((np_, np_), 1.0 * dt / dx), np.full((np_ + 1, np_ - 1), 2.0 * dt / dy)]
    else:
        psi = 2 + (np.sin(2 * np.pi * ax[:, None, None] * dx) * np.sin(2 * np.pi * ax[None, :, None] * dy)
                   * np.sin(2 * np.pi * ax[None, None, :] * dz))
        gc = [np.full((np_, np_, np_), 1.0 * dt / dx), np.full((np_ + 1, np_ - 1, np_), 2.0 * dt / dy),
              np.full((np_ + 1, np_, np_ - 1), 3.0 * dt / dy)]
    return dict(grid=(np_,) * nd, bcs=[(0, 0)] * nd, n_iters=2, nt=nt, psi=psi, gc=gc, time=time)


CONVERGENCE_SIZES = {2: (33, 65), 3: (17, 33)}


def l2_error(out, case):
    n = case["psi"].size
    return np.sqrt(np.sum((out - case["psi"]) ** 2) / n) / case["time"]


def basic_example_case():
    """tests/paper_2015_GMD/0_basic_example/basic_example.cpp: 1D, 101 open
    cells, dx = 0.1, the default options (iga | fct, n_iters = 2), a Witch
    of Agnesi -0.5 + 1 / ((dx (i - 50))^2 + 1) under Courant 0.5, output
    every 20 steps up to 100; here as 101 x 4 with no flow along y."""
    nx, ny, dx = 101, 4, 0.1
    i = np.arange(nx)
    psi = np.repeat((-.5 + 1 / ((dx * (i - (nx - 1) / 2.)) ** 2 + 1))[:, None], ny, axis=1)
    gc = [np.full((nx, ny), .5), np.zeros((nx + 1, ny - 1))]
    return dict(grid=(nx, ny), bcs=[(1, 1), (1, 1)], n_iters=2, psi=psi, gc=gc, outfreq=20, nt=100)


def basic_example_records(golden_dir):
    import os
    path = os.path.join(golden_dir, "basic_example", "records.txt")
    return [np.array([float(v) for v in line.split()]) for line in open(path)]


# option groups of 1_advscheme_opts' log, in order; the tot ones are not
# built by this backend
ADVSCHEME_GROUPS = ["abs", "iga", "iga|tot", "iga|fct", "iga|tot|fct"]


def advscheme_opts_case():
    """tests/paper_2015_GMD/1_advscheme_opts/advscheme_opts.cpp: 1D, 601
    cyclic cells, two equations (2 / 4 and -1 / +1 steps over cells
    76..124), Courant -0.75, n_iters = 2, 1200 steps; here as 601 x 4 with
    no flow along y."""
    nx, ny = 601, 4
    i = np.arange(nx)
    out = (i <= 75) | (i >= 125)
    psi = [np.repeat(np.where(out, 2.0, 4.0)[:, None], ny, axis=1),
           np.repeat(np.where(out, -1.0, 1.0)[:, None], ny, axis=1)]
    gc = [np.full((nx, ny), -.75), np.zeros((nx + 1, ny - 1))]
    return dict(grid=(nx, ny), bcs=[(0, 0), (1, 1)], n_iters=2, psi=psi, gc=gc, nt=1200)


def advscheme_opts_finals(golden_dir, group):
    """final fields (equation 0, equation 1) of option group `group` from the
    log: of its four blocks, the two that are not the initial fields"""
    import os
    path = os.path.join(golden_dir, "advscheme_opts", "records.txt")
    blocks = [np.array([float(v) for v in line.split()]) for line in open(path)]
    g = blocks[4 * group:4 * group + 4]
    c = advscheme_opts_case()
    init = [p[:, 0] for p in c["psi"]]
    finals = [b for b in g if not any(np.array_equal(b, x) for x in init)]
    assert len(finals) == 2
    finals.sort(key=lambda b: -b.mean())        # equation 0 (2 / 4) first
    return finals
