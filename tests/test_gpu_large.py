"""A grid of more than 2^31 cells (2048 x 1040 x 1040, FP32, FCT n_iters=3,
cyclic): every index path (interior tiles, boundary shell, halo fills,
copies, the mass reduction) must stay 64-bit clean.  The field varies only
along x and the advector is a uniform x-Courant number, so every (y, z)
column must evolve like the same problem on a 2048 x 16 x 16 grid; the
mass of the unique cells is conserved and FCT creates no new extrema."""
import math

import numpy as np
import pytest
import torch

import libmpdataxx_amd as M

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
NX = 2048
C = 0.25
STEPS = 4


def profile_x(nx):
    """A top hat and a Gaussian on 0.2, periodic with period nx-1 (cell
    nx-1 duplicates cell 0)."""
    x = torch.arange(nx, dtype=torch.float64, device=DEV)
    per = nx - 1

    def pdist(c):
        d = torch.remainder(x - c, per)
        return torch.minimum(d, per - d)
    f = 0.2 + torch.where(pdist(per / 3) <= 100, 1.0, 0.0) + 0.5 * torch.exp(-pdist(1200.0) ** 2 / 800)
    return f.to(torch.float32)


def run(grid):
    s = M.Solver(grid, opts=M.FCT, n_iters=3, dtype=np.float32, bcs=[(M.CYCLIC, M.CYCLIC)] * 3)
    prof = profile_x(grid[0])
    psi = prof[:, None, None].expand(grid).contiguous()
    s.set_advectee_device(psi)
    del psi
    for d in range(3):
        g = torch.full(s.advector_shape(d), C if d == 0 else 0.0, dtype=torch.float32, device=DEV)
        s.set_advector_device(g, d)
        del g
    torch.cuda.empty_cache()
    return s, prof


def test_more_than_2_31_cells():
    big = (NX, 1040, 1040)
    assert math.prod(big) > 2 ** 31
    print(f"large grid {big} ({math.prod(big) / 2 ** 31:.2f} x 2^31 cells) FCT n_iters=3 {STEPS} steps", flush=True)
    s, prof = run(big)
    m0 = s.mass()
    s.advance(STEPS)
    m1 = s.mass()
    out = torch.empty(big, dtype=torch.float32, device=DEV)
    s.get_advectee_device(out)
    s.close()
    torch.cuda.empty_cache()
    small = (NX, 16, 16)
    r, _ = run(small)
    r.advance(STEPS)
    ref = torch.empty(small, dtype=torch.float32, device=DEV)
    r.get_advectee_device(ref)
    r.close()
    col = ref[:, 8, 8].double()
    # the field moved (x-Courant 0.25) and every column equals the small run's
    assert float((col - prof.double()).abs().max()) > 0.1
    for y, z in [(0, 0), (1, 1), (2, 517), (520, 520), (1037, 3), (1038, 1038), (1039, 1039), (700, 1039)]:
        c = out[:, y, z].double()
        assert float((c - col).abs().max()) <= 1e-6 * float(col.abs().max()), (y, z)
    # no new extrema; conservation over the unique cells of a column; the
    # FP64 Kahan sum over all > 2^31 cells (duplicated edge cells included)
    # equals the column sums times the number of columns
    assert float(out.min()) >= float(prof.min()) * (1 - 2 ** -20)
    assert float(out.max()) <= float(prof.max()) * (1 + 2 ** -20)
    u0, u1 = float(prof[:-1].double().sum()), float(col[:-1].sum())
    assert abs(u1 - u0) <= 1e-6 * u0
    ncol = big[1] * big[2]
    assert abs(m0 - ncol * float(prof.double().sum())) <= 1e-9 * m0
    assert abs(m1 - ncol * float(col.sum())) <= 1e-6 * m1
    del out
    torch.cuda.empty_cache()
