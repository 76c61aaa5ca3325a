"""Helpers shared by the GPU parity tests (not collected by pytest)."""
import numpy as np

import oracle as O
from libmpdataxx_amd import workloads as W


def setup_case(grid, bcs, seed, field="random"):
    nd = len(grid)
    per = tuple(b[0] == O.CYCLIC for b in bcs)
    gc = W.streamfunction_advector(grid, seed, periodic=per, max_courant=0.45)
    if field == "random":
        psi = W.random_field(grid, seed + 1000)
    else:
        psi = field
    return psi, gc


def run_oracle(grid, bcs, opts, n_iters, psi, gc, nt, n_eqns=1, **kw):
    o = O.Oracle(grid, opts=opts, n_iters=n_iters, bcs=bcs, n_eqns=n_eqns,
                 max_abs_div_eps=kw.pop("max_abs_div_eps", 1e300), **kw)
    psis = psi if isinstance(psi, list) else [psi]
    for e, p in enumerate(psis):
        o.set_advectee(p, e)
    for d in range(len(grid)):
        o.set_advector(gc[d], d)
    o.advance(nt)
    return [o.advectee(e) for e in range(len(psis))]


def run_gpu(grid, bcs, opts, n_iters, psi, gc, nt, n_eqns=1, dtype=np.float64, **kw):
    import libmpdataxx_amd as M
    s = M.Solver(grid, opts=opts, n_iters=n_iters, bcs=bcs, n_eqns=n_eqns, dtype=dtype,
                 max_abs_div_eps=kw.pop("max_abs_div_eps", 1e300), **kw)
    psis = psi if isinstance(psi, list) else [psi]
    for e, p in enumerate(psis):
        s.set_advectee(p, e)
    for d in range(len(grid)):
        s.set_advector(gc[d], d)
    s.advance(nt)
    out = [s.advectee(e).astype(np.float64) for e in range(len(psis))]
    s.close()
    return out


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
