"""Device-side (torch) generators of the large benchmark workloads, so that
512^3 inputs are created in HBM without host round trips.  Same fields as
the numpy generators of workloads.py where both exist."""
from __future__ import annotations

import math

import torch


def rotation_sphere_device(n, dtype, dev, max_courant=0.5):
    """workloads.rotation_sphere_3d generated on the GPU: open N^3 solid-body
    rotation about z moving a discontinuous sphere (radius 40 n / 256, centre
    offset 48 n / 256 along x)."""
    c = (n - 1) / 2.0
    w = max_courant / c
    radius, offset = 40.0 * n / 256, 48.0 * n / 256
    i = torch.arange(n, dtype=torch.float64, device=dev)
    r2 = ((i[:, None, None] - (c + offset)) ** 2 + (i[None, :, None] - c) ** 2
          + (i[None, None, :] - c) ** 2)
    psi = torch.where(r2 <= radius ** 2, 1.0, 0.0).to(dtype)
    del r2
    gx = (-w * (i - c)).to(dtype)[None, :, None].expand(n, n, n).contiguous()
    ia = torch.arange(-1, n, dtype=torch.float64, device=dev)
    gy = (w * (ia - c)).to(dtype)[:, None, None].expand(n + 1, n - 1, n).contiguous()
    gz = torch.zeros((n + 1, n, n - 1), dtype=dtype, device=dev)
    return psi, [gx, gy, gz]


def multi_tracer_fields(shape, dev):
    """The eight FP32 tracers of the multi-tracer workload on a cyclic grid
    of period n-1 (cell n-1 duplicates cell 0): 0-3 lognormal exp(g/2) with g
    a sum of 8 random Fourier modes, 4-6 slotted cylinders along z, 7 the
    constant 1."""
    n0, n1, n2 = shape
    out = []
    # cell n-1 is evaluated as cell 0, so the duplicated edge repeats bit for bit
    ix = [torch.arange(n, dtype=torch.float64, device=dev) % (n - 1) for n in shape]
    X, Y, Z = ix[0][:, None, None], ix[1][None, :, None], ix[2][None, None, :]
    for t in range(8):
        g = torch.Generator().manual_seed(4000 + t)
        if t < 4:
            f = torch.zeros(shape, dtype=torch.float64, device=dev)
            for _ in range(8):
                k = torch.randint(-4, 5, (3,), generator=g).tolist()
                ph = float(torch.rand(1, generator=g)) * 2 * math.pi
                f += 0.5 * torch.cos(2 * math.pi * (k[0] * X / (n0 - 1) + k[1] * Y / (n1 - 1)
                                                    + k[2] * Z / (n2 - 1)) + ph)
            f = torch.exp(0.5 * f)
        elif t < 7:
            c = (torch.rand(2, generator=g) * 0.5 + 0.25).tolist()
            cx, cy, r = c[0] * (n0 - 1), c[1] * (n1 - 1), 0.15 * (n0 - 1)
            inside = ((X - cx) ** 2 + (Y - cy) ** 2 <= r * r) & ~((abs(X - cx) <= r / 6) & (Y - cy >= -r / 3))
            f = torch.where(inside, 1.0, 0.1).expand(shape).clone()
        else:
            f = torch.ones(shape, dtype=torch.float64, device=dev)
        out.append(f.to(torch.float32).contiguous())
    return out
