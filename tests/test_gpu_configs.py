"""The benchmark configurations as parity cases, at their full sizes.

C1  2D MPDATA n_iters=2, FP64, 256^2 cyclic, uniform Courant (0.25, 0.25),
    Gaussian cone on 0.5, 200 steps: GPU against the oracle at full size.
C2  3D MPDATA-FCT n_iters=3, FP64, 256^3 open, solid-body rotation, sphere
    of 1 on 0, 100 steps: FCT bounds, mass, FP32 accuracy at full size; the
    oracle at 64^3.
C3  the same at 512^3 in FP32 against FP64 on the GPU, 100 steps: L2 error,
    mass error, bounds (the oracle pins FP64 at smaller sizes).
C4  8 tracers, 512x512x128 cyclic, on-device reversing deformational flow,
    FCT n_iters=3, FP32, 200 steps: constant tracer, per-tracer mass over the
    unique cells, bounds (the oracle pins the generator at smaller sizes).
"""
import numpy as np
import pytest
import torch

import libmpdataxx_amd as M
import oracle as O
from libmpdataxx_amd import workloads as W
from libmpdataxx_amd.workloads_torch import multi_tracer_fields, rotation_sphere_device

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def test_c1_gaussian_2d_cyclic_matches_oracle():
    print("C1 256^2 cyclic FP64 200 steps", flush=True)
    psi, gc = W.gaussian_2d_cyclic(256)
    bcs = [(M.CYCLIC, M.CYCLIC)] * 2
    s = M.Solver((256, 256), opts=0, n_iters=2, bcs=bcs, dtype=np.float64)
    o = O.Oracle((256, 256), opts=0, n_iters=2, bcs=bcs)
    for x in (s, o):
        x.set_advectee(psi)
        for d in range(2):
            x.set_advector(gc[d], d)
    s.advance(200)
    o.advance(200)
    assert s.z_tiles == 3   # all-cyclic and repeating: the 2D pass kernel alone
    got, want = s.advectee(), o.advectee()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # cyclic with period n-1: the duplicated edge stays a copy, the mass of
    # the unique cells is conserved
    np.testing.assert_array_equal(got[0, :], got[-1, :])
    np.testing.assert_array_equal(got[:, 0], got[:, -1])
    m0, m1 = psi[:-1, :-1].sum(), got[:-1, :-1].sum()
    assert abs(m1 - m0) <= 1e-13 * m0
    s.close()


@pytest.mark.parametrize("n,steps", [(64, 20), (256, 10)])
def test_c2_rotation_sphere_matches_oracle(n, steps):
    """C2 against the strict FP64 oracle, at 64^3 and at its full 256^3
    (fields in [0, 1]: absolute agreement 1e-12)."""
    print(f"C2 at {n}^3 FP64 {steps} steps vs oracle", flush=True)
    psi, gc = W.rotation_sphere_3d(n)
    s = M.Solver((n,) * 3, opts=M.FCT, n_iters=3, dtype=np.float64)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(gc[d], d)
    s.advance(steps)
    got = s.advectee()
    s.close()
    o = O.Oracle((n,) * 3, opts=O.FCT, n_iters=3)
    o.set_advectee(psi)
    for d in range(3):
        o.set_advector(gc[d], d)
    del psi, gc
    o.advance(steps)
    want = o.advectee()
    del o
    assert np.abs(got - want).max() <= 1e-12


def _sphere(n, dtype, steps):
    s = M.Solver((n,) * 3, opts=M.FCT, n_iters=3, dtype=np.float32 if dtype == torch.float32 else np.float64)
    psi, gc = rotation_sphere_device(n, dtype, DEV)
    s.set_advectee_device(psi)
    for d in range(3):
        s.set_advector_device(gc[d], d)
    del psi, gc
    m0 = s.mass()
    s.advance(steps)
    out = torch.empty((n,) * 3, dtype=dtype, device=DEV)
    s.get_advectee_device(out)
    return s, out, m0


@pytest.mark.parametrize("n", [256, 512])
def test_c2_c3_rotation_sphere_full_size(n):
    """FCT creates no new extrema (the field starts in {0, 1}); nothing
    reaches the open boundaries, so mass is conserved; FP32 stays within the
    accuracy bounds against the FP64 run (L2 <= 1e-5 relative, mass <= 1e-6)."""
    steps = 100
    print(f"C2/C3 {n}^3 FCT n_iters=3 {steps} steps FP64 and FP32", flush=True)
    s64, a64, m64_0 = _sphere(n, torch.float64, steps)
    assert float(a64.min()) >= 0.0 and float(a64.max()) <= 1.0
    assert abs(s64.mass() - m64_0) <= 1e-12 * m64_0
    s64.close()
    s32, a32, m32_0 = _sphere(n, torch.float32, steps)
    assert float(a32.min()) >= 0.0 and float(a32.max()) <= 1.0
    assert abs(s32.mass() - m32_0) <= 1e-6 * m32_0
    d = a32.double() - a64
    l2 = float(torch.sqrt((d ** 2).sum() / (a64 ** 2).sum()))
    print(f"  L2 rel FP32 vs FP64 = {l2:.3e}", flush=True)
    assert l2 <= 1e-5
    s32.close()


def test_c4_multi_tracer_full_size():
    shape, steps = (512, 512, 128), 200
    print(f"C4 {shape} 8 tracers {steps} steps", flush=True)
    s = M.Solver(shape, opts=M.FCT, n_iters=3, n_eqns=8, dtype=np.float32, dt=1.0,
                 bcs=[(M.CYCLIC, M.CYCLIC)] * 3)
    fields = multi_tracer_fields(shape, DEV)
    for e, f in enumerate(fields):
        s.set_advectee_device(f, e)
    for d in range(3):
        s.set_advector_device(torch.zeros(s.advector_shape(d), dtype=torch.float32, device=DEV), d)
    s.set_advector_generator(M.GC_DEFORM, 200.0, 0.4)
    lo = [float(f.min()) for f in fields]
    hi = [float(f.max()) for f in fields]
    m0 = [float(f[:-1, :-1, :-1].double().sum()) for f in fields]
    del fields
    s.advance(steps)
    assert s.z_tiles == 3   # repeating planes everywhere: the interior tiles take every cyclic bound
    out = torch.empty(shape, dtype=torch.float32, device=DEV)
    for e in range(8):
        s.get_advectee_device(out, e)
        u = out[:-1, :-1, :-1].double()
        m = float(u.sum())
        assert abs(m - m0[e]) <= 1e-6 * m0[e], (e, m, m0[e])
        assert float(out.min()) >= lo[e] * (1 - 2 ** -20) and float(out.max()) <= hi[e] * (1 + 2 ** -20), e
        # cyclic: the duplicated edge cells stay copies of cell 0
        assert torch.equal(out[0], out[-1]) and torch.equal(out[:, 0], out[:, -1]), e
        assert torch.equal(out[:, :, 0], out[:, :, -1]), e
    s.get_advectee_device(out, 7)
    assert float((out - 1).abs().max()) <= 4 * 2.0 ** -23
    s.close()
