"""Initial conditions and advector fields for the standard MPDATA workloads.

Pure numpy, FP64.  Every generator returns ``(psi, [advector(d) ...])`` laid
out like the reference's ``advectee()`` / ``advector(d)`` views:
``advector(d)`` spans faces -1..n_d-2 along d (index i <-> face i+1/2) and,
for 3D components d > 0, x = -1..n0-1.  Arrays are C-order ``[x][y][z]``.

Test problems follow the reference's own setups:
  * revolving sphere 3D  -- tests/paper_2015_GMD/4_revolving_sphere_3d/
                            revolving_sphere_3d.cpp:29-97
  * rotating cone 2D     -- tests/paper_2015_GMD/3_rotating_cone_2d/
                            rotating_cone_2d.cpp:32-140
Benchmark workloads (no reference numbers exist for these shapes):
  * gaussian_2d_cyclic   -- 256^2 cyclic, uniform Courant (0.25, 0.25)
  * rotation_sphere_3d   -- N^3 open, solid-body rotation about z
  * deform_multitracer   -- 512x512x128 cyclic, 8 tracers, reversing flow
Random numbers come from numpy's PCG64 with fixed seeds.
"""
from __future__ import annotations

import math

import numpy as np


def face_index(n, lo=-1):
    return np.arange(lo, lo + n)


# ---------------------------------------------------------------- reference

def revolving_sphere_3d(nx=59):
    """revolving_sphere_3d.cpp:29-97 (FP64, open boundaries)."""
    dt = 0.018 * 2 * math.pi
    L = 100.0
    dx = L / (nx - 1)
    dy = dz = dx
    h = 4.0
    r = 15.0
    d = 25 / math.sqrt(3)
    x0, y0, z0 = 50 - d, 50 + d, 50 + d
    i = np.arange(nx, dtype=np.float64)[:, None, None]
    j = np.arange(nx, dtype=np.float64)[None, :, None]
    k = np.arange(nx, dtype=np.float64)[None, None, :]
    tmp = (i * dx - x0) ** 2 + (j * dx - y0) ** 2 + (k * dx - z0) ** 2
    psi = np.where(tmp - r ** 2 <= 0, h, 0.0)
    omega = 0.1
    xc = yc = zc = 50.0
    c = omega / math.sqrt(3)
    ia = np.arange(-1, nx, dtype=np.float64)            # x = -1..n-1
    gx = np.empty((nx, nx, nx))
    jj = np.arange(nx, dtype=np.float64)[None, :, None]
    kk = np.arange(nx, dtype=np.float64)[None, None, :]
    gx[:] = c * (-(jj * dy - yc) + (kk * dz - zc)) * dt / dx
    gy = np.empty((nx + 1, nx - 1, nx))
    gy[:] = c * ((ia[:, None, None] * dx - xc) - (kk * dz - zc)) * dt / dy
    gz = np.empty((nx + 1, nx, nx - 1))
    gz[:] = c * (-(ia[:, None, None] * dx - xc) + (jj * dy - yc)) * dt / dz
    return psi, [gx, gy, gz], dt


def revolving_sphere_stats(ic, fc, timestep, dt):
    """revolving_sphere_stats.hpp:46-65"""
    n = ic.shape[0]
    diff = fc - ic
    return {
        "max": float(fc.max()),
        "min": float(fc.min()),
        "Linf": float(np.abs(diff).max()),
        "L2": float(1.0 / (timestep * dt) * math.sqrt(1.0 / n ** 3 * float((diff ** 2).sum()))),
        "L1": float(1.0 / (timestep * dt) * (1.0 / n ** 3 * float(np.abs(diff).sum()))),
    }


def rotating_cone_2d(n=101):
    """rotating_cone_2d.cpp:32-140 (FP64, open boundaries)."""
    dt, dx, dy, omega, h, h0 = 0.1, 1.0, 1.0, 0.1, 4.0, 1.0
    r = 15.0 * dx
    x0, y0 = 50 * dx, 75 * dy
    xc = 0.5 * (n - 1) * dx
    yc = 0.5 * (n - 1) * dy
    i = np.arange(n, dtype=np.float64)[:, None]
    j = np.arange(n, dtype=np.float64)[None, :]
    tmp = (i * dx - x0) ** 2 + (j * dy - y0) ** 2
    psi = h0 + np.where(tmp - r ** 2 <= 0, h - np.sqrt(tmp / (r / h) ** 2), 0.0)
    gx = np.empty((n, n))
    gx[:] = omega * (j * dy - yc) * dt / dx
    ia = np.arange(-1, n, dtype=np.float64)[:, None]
    gy = np.empty((n + 1, n - 1))
    gy[:] = -omega * (ia * dx - xc) * dt / dy
    return psi, [gx, gy], dt


def rotating_cone_stats(ic, fc, timestep, dt):
    """rotating_cone_stats.hpp:54-69"""
    err2 = (ic - fc) ** 2
    return {
        "min": float(fc.min()),
        "max": float(fc.max()),
        "rms": float(math.sqrt(err2.sum() / ic.shape[0] / ic.shape[1]) / timestep / dt),
    }


# ---------------------------------------------------------------- benchmarks

def gaussian_2d_cyclic(n=256, courant=(0.25, 0.25), amp=1.0, sigma=16.0, bg=0.5):
    """2D cyclic Gaussian cone on a background, uniform advector.

    The cyclic domain repeats with period n-1 (cells 0 and n-1 coincide), so
    the distance is periodic with period n-1 and psi[0] == psi[n-1]."""
    per = n - 1
    c = n // 2
    idx = np.arange(n, dtype=np.float64)
    dist = np.abs(idx - c)
    dist = np.minimum(dist, per - dist)
    r2 = dist[:, None] ** 2 + dist[None, :] ** 2
    psi = bg + amp * np.exp(-r2 / (2 * sigma ** 2))
    gx = np.full((n, n), courant[0])
    gy = np.full((n + 1, n - 1), courant[1])
    return psi, [gx, gy]


def rotation_sphere_3d(n=256, radius=None, offset=None, max_courant=0.5):
    """Open-boundary N^3 solid-body rotation about z with a discontinuous
    sphere (1 inside, 0 outside).  GC_x = -w (j - c), GC_y = w (i - c),
    GC_z = 0 with c = (n-1)/2 and w = max_courant / c, so the discrete
    divergence is exactly zero."""
    if radius is None:
        radius = 40.0 * n / 256
    if offset is None:
        offset = 48.0 * n / 256
    c = (n - 1) / 2.0
    w = max_courant / c
    i = np.arange(n, dtype=np.float64)
    r2 = ((i[:, None, None] - (c + offset)) ** 2 + (i[None, :, None] - c) ** 2
          + (i[None, None, :] - c) ** 2)
    psi = np.where(r2 <= radius ** 2, 1.0, 0.0)
    gx = np.empty((n, n, n))
    gx[:] = (-w * (i - c))[None, :, None]
    ia = np.arange(-1, n, dtype=np.float64)
    gy = np.empty((n + 1, n - 1, n))
    gy[:] = (w * (ia - c))[:, None, None]
    gz = np.zeros((n + 1, n, n - 1))
    return psi, [gx, gy, gz]


def random_field(shape, seed, lo=0.1, hi=1.0):
    """Seeded positive field for parity tests."""
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=shape)


def streamfunction_advector(grid, seed, max_courant=0.5, n_modes=4, periodic=True,
                            quantum=2.0 ** -20, exact_period=False):
    """Discretely divergence-free random advector from a streamfunction.

    The streamfunction lives on cell corners and is quantised to a dyadic
    grid so that every advector value and every divergence sum is exact.
    For periodic domains the corner field has period n-1 so the advector is
    consistent across the duplicated edge cell (bit for bit with
    exact_period, which evaluates a corner at its index reduced modulo the
    period).  3D: horizontal (x, y) flow per z-level plus a w built from a
    second streamfunction in (x, z)."""
    rng = np.random.default_rng(seed)
    nd = len(grid)
    if isinstance(periodic, bool):
        periodic = (periodic,) * nd

    def corner_field(nx, ny, perx, pery):
        # corners at x = i+1/2 for i = -2..nx-1, same in y (covers all faces)
        xs = np.arange(-2, nx) + 0.5
        ys = np.arange(-2, ny) + 0.5
        px = (nx - 1) if perx else 2 * nx
        py = (ny - 1) if pery else 2 * ny
        if exact_period:
            xs = np.mod(np.arange(-2, nx), px) + 0.5
            ys = np.mod(np.arange(-2, ny), py) + 0.5
        f = np.zeros((xs.size, ys.size))
        for _ in range(n_modes):
            kx = rng.integers(1, 3)
            ky = rng.integers(1, 3)
            ph1, ph2 = rng.uniform(0, 2 * np.pi, 2)
            a = rng.normal()
            f += a * np.sin(2 * np.pi * kx * xs[:, None] / px + ph1) * np.sin(2 * np.pi * ky * ys[None, :] / py + ph2)
        return f

    if nd == 2:
        nx, ny = grid
        f = corner_field(nx, ny, periodic[0], periodic[1])
        # GC_x(face i+1/2, cell j) = f(i+1/2, j+1/2) - f(i+1/2, j-1/2)
        gx_full = f[:, 1:] - f[:, :-1]          # x corners -2..nx-1 -> faces; y cells -1..ny-1
        gy_full = -(f[1:, :] - f[:-1, :])       # x cells -1..nx-1; y faces -2..ny-1
        scale = max_courant / max(np.abs(gx_full).max(), np.abs(gy_full).max())
        f = np.round(f * scale / quantum) * quantum
        gx_full = f[:, 1:] - f[:, :-1]
        gy_full = -(f[1:, :] - f[:-1, :])
        # views: gx faces -1..nx-2 (corner index +1), y cells 0..ny-1 (index +1)
        gx = gx_full[1:nx + 1, 1:ny + 1]
        gy = gy_full[0:nx + 1, 2:ny + 1]
        return [np.ascontiguousarray(gx), np.ascontiguousarray(gy)]
    nx, ny, nz = grid
    f = corner_field(nx, ny, periodic[0], periodic[1])
    g = corner_field(nx, nz, periodic[0], periodic[2])
    gx_full = f[:, 1:] - f[:, :-1]
    gy_full = -(f[1:, :] - f[:-1, :])
    hx_full = g[:, 1:] - g[:, :-1]           # x faces, z cells
    hz_full = -(g[1:, :] - g[:-1, :])        # x cells, z faces
    scale = max_courant / max(np.abs(gx_full).max() + np.abs(hx_full).max(),
                              np.abs(gy_full).max(), np.abs(hz_full).max())
    f = np.round(f * scale / quantum) * quantum
    g = np.round(g * scale / quantum) * quantum
    gx_full = f[:, 1:] - f[:, :-1]
    gy_full = -(f[1:, :] - f[:-1, :])
    hx_full = g[:, 1:] - g[:, :-1]
    hz_full = -(g[1:, :] - g[:-1, :])
    gx = np.empty((nx, ny, nz))
    gx[:] = gx_full[1:nx + 1, 1:ny + 1][:, :, None] + hx_full[1:nx + 1, 1:nz + 1][:, None, :]
    gy = np.empty((nx + 1, ny - 1, nz))
    gy[:] = gy_full[0:nx + 1, 2:ny + 1][:, :, None]
    gz = np.empty((nx + 1, ny, nz - 1))
    gz[:] = hz_full[0:nx + 1, 2:nz + 1][:, None, :]
    return [gx, gy, gz]


def coupled_harmosc(n=1001, ny=4):
    """Coupled harmonic oscillator advected at Courant 0.5 on a cyclic grid
    (tests/paper_2015_GMD/6_coupled_harmosc/coupled_harmosc.cpp:40-70): psi a
    cosine bump on cells 51..149, phi = 0, omega = 2 pi / 400 (dt = 1).  The
    reference's 1D problem is embedded in a y-uniform 2D grid with zero
    y-advector (the 2D formulae reduce to the 1D ones on it).  Returns
    (psi, phi, gc, omega)."""
    i = np.arange(n)
    psi1 = np.where((i <= 50) | (i >= 150), 0.0, 0.5 * (1 + np.cos(2 * np.pi * i / 100)))
    psi = np.repeat(psi1[:, None], ny, 1)
    phi = np.zeros((n, ny))
    gc = [np.full((n, ny), 0.5), np.zeros((n + 1, ny - 1))]
    return psi, phi, gc, 2 * np.pi / 400


def coupled_harmosc_rhs(psi, phi, dt, at, omega):
    """update_rhs of tests/paper_2015_GMD/6_coupled_harmosc/coupled_harmosc.hpp:
    26-56: explicit R^n (at = 0), implicit R^{n+1} (at = 1)."""
    if at == 0:
        return omega * phi, -omega * psi
    d = 1 + (dt * omega) ** 2
    return ((psi + dt * omega * phi) / d - psi) / dt, ((phi - dt * omega * psi) / d - phi) / dt


def coupled_harmosc_stats(psi, phi, timestep, dt=1.0, vel=0.5):
    """min / max with first index and the rms error against the advected
    bump (coupled_harmosc_stats.hpp:24-72), on 1D rows."""
    n = psi.size
    t = timestep * dt
    i = np.arange(n)
    sol = np.where((i <= 50 + vel * t) | (i >= 150 + vel * t), 0.0, 0.5 * (1 + np.cos(2 * np.pi * i / 100.)))
    tmp = (psi ** 2 + phi ** 2 - sol ** 2) ** 2
    return {"timestep": timestep,
            "min_psi": (psi.min(), int(psi.argmin())), "max_psi": (psi.max(), int(psi.argmax())),
            "min_phi": (phi.min(), int(phi.argmin())), "max_phi": (phi.max(), int(phi.argmax())),
            "rms": np.sqrt((tmp.sum() - tmp[-1]) / (n - 1)) / timestep}


def parse_coupled_harmosc_stats(text):
    """The refdata stats.txt of the coupled oscillator as a list of dicts."""
    import re
    out, cur = [], None
    for line in text.splitlines():
        m = re.match(r"timestep = (\d+)", line)
        if m:
            cur = {"timestep": int(m.group(1))}
            out.append(cur)
            continue
        m = re.match(r"(min|max)\((psi|phi)\) =\s*(\S+) @ index \((\d+)\)", line)
        if m and cur is not None:
            cur[f"{m.group(1)}_{m.group(2)}"] = (float(m.group(3)), int(m.group(4)))
            continue
        m = re.match(r"rms = (\S+)", line)
        if m and cur is not None:
            cur["rms"] = float(m.group(1))
    return out


# ---------------------------------------------------------------- ensembles

class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto & Nishimura 2004), as
    std::mt19937_64, so draws match a C++ host's."""

    def __init__(self, seed=5489):
        m = (1 << 64) - 1
        self.mt = [0] * 312
        self.mt[0] = seed & m
        for i in range(1, 312):
            x = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (x ^ (x >> 62)) + i) & m
        self.i = 312

    def __call__(self):
        m = (1 << 64) - 1
        if self.i >= 312:
            mt = self.mt
            for k in range(312):
                y = (mt[k] & 0xFFFFFFFF80000000) | (mt[(k + 1) % 312] & 0x7FFFFFFF)
                mt[k] = mt[(k + 156) % 312] ^ (y >> 1) ^ (0xB5026F5AA96619E9 if y & 1 else 0)
            self.i = 0
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & m

    def uniform_int(self, lo, hi):
        """lo..hi inclusive (modulo reduction, hand-written)."""
        return lo + self() % (hi - lo + 1)


def _periodic_dist(n, c):
    per = n - 1
    d = np.abs(np.arange(n, dtype=np.float64) - c) % per
    return np.minimum(d, per - d)


def level_streamfunction_advector(grid, seed, max_courant=0.5, n_modes=8, quantum=2.0 ** -20):
    """3D cyclic advector with horizontal flow only (w = 0): a random
    periodic 2D streamfunction of n_modes Fourier modes whose amplitude varies
    with the level, quantised per level so that every level's discrete
    divergence is exactly zero."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = grid
    xs = np.arange(-2, nx) + 0.5
    ys = np.arange(-2, ny) + 0.5
    f = np.zeros((xs.size, ys.size))
    for _ in range(n_modes):
        kx, ky = rng.integers(1, 4, 2)
        ph1, ph2 = rng.uniform(0, 2 * np.pi, 2)
        f += rng.normal() * np.sin(2 * np.pi * kx * xs[:, None] / (nx - 1) + ph1) \
            * np.sin(2 * np.pi * ky * ys[None, :] / (ny - 1) + ph2)
    phz = rng.uniform(0, 2 * np.pi)
    # period nz-1, exactly: level nz-1 is evaluated as level 0
    lev = 1.0 + 0.5 * np.sin(2 * np.pi * (np.arange(nz) % (nz - 1)) / (nz - 1) + phz)
    gx0 = f[:, 1:] - f[:, :-1]
    gy0 = -(f[1:, :] - f[:-1, :])
    scale = max_courant / (max(np.abs(gx0).max(), np.abs(gy0).max()) * lev.max())
    fl = np.round(f[:, :, None] * (scale * lev)[None, None, :] / quantum) * quantum
    gx_full = fl[:, 1:, :] - fl[:, :-1, :]
    gy_full = -(fl[1:, :, :] - fl[:-1, :, :])
    gx = np.ascontiguousarray(gx_full[1:nx + 1, 1:ny + 1, :])
    gy = np.ascontiguousarray(gy_full[0:nx + 1, 2:ny + 1, :])
    gz = np.zeros((nx + 1, ny, nz - 1))
    return [gx, gy, gz]


def ensemble_domains(n2=32, n3=32, seed=5000, max_courant=0.5):
    """The uneven ensemble: n2 2D domains with sides drawn from 64..512 and
    n3 3D domains with sides from 32..128 (std::mt19937_64(seed)); domain d
    (generator seed + d) draws n_iters in {2, 3, 4} and FCT on/off, and gets a
    divergence-free random 8-mode streamfunction advector scaled to
    max_courant and psi = U(0.1, 1) + 3 Gaussian blobs, all cyclic (period
    n-1 per dimension, the last cell duplicating the first).  Returns a list
    of dicts: grid, n_iters, fct, psi, gc."""
    g = MT19937_64(seed)
    sizes = [tuple(g.uniform_int(64, 512) for _ in range(2)) for _ in range(n2)]
    sizes += [tuple(g.uniform_int(32, 128) for _ in range(3)) for _ in range(n3)]
    out = []
    for d, grid in enumerate(sizes):
        r = MT19937_64(seed + d)
        n_iters = r.uniform_int(2, 4)
        fct = bool(r() & 1)
        field_seed = r()
        rng = np.random.default_rng(field_seed)
        nd = len(grid)
        unique = tuple(n - 1 for n in grid)
        psi = np.empty(grid)
        core = rng.uniform(0.1, 1.0, size=unique)
        for _ in range(3):
            c = [rng.uniform(0, n - 1) for n in grid]
            sig = rng.uniform(2.0, 6.0)
            amp = rng.uniform(0.5, 2.0)
            r2 = sum(_periodic_dist(n, cc)[tuple(slice(None) if q == k else None for q in range(nd))] ** 2
                     for k, (n, cc) in enumerate(zip(grid, c)))
            core = core + amp * np.exp(-r2[tuple(slice(0, u) for u in unique)] / (2 * sig * sig))
        psi[tuple(slice(0, u) for u in unique)] = core
        # duplicate the first cell of every dimension into the last
        for k in range(nd):
            src = [slice(None)] * nd
            dst = [slice(None)] * nd
            src[k], dst[k] = 0, grid[k] - 1
            psi[tuple(dst)] = psi[tuple(src)]
        adv_seed = int(field_seed % (1 << 63))
        if nd == 2:
            gc = streamfunction_advector(grid, adv_seed, max_courant=max_courant, n_modes=8, periodic=True,
                                         exact_period=True)
        else:
            gc = level_streamfunction_advector(grid, adv_seed, max_courant=max_courant)
        out.append({"grid": grid, "n_iters": n_iters, "fct": fct, "psi": psi, "gc": gc})
    return out
