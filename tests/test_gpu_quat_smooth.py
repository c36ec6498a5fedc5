"""The one-pass orientation red-black sweep (ampe_quat_smooth with fused = 1)
against the red launch + black launch form (fused = 0): identical values,
bit for bit, after 1-3 sweeps, in 2-D and 3-D, on one periodic patch, on
uneven patches and on a non-periodic domain (patch outer layers exercise the
shell pass)."""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu


def _boxes(shape, split):
    ndim = len(shape)
    ext = list(shape[::-1]) + [1] * (3 - ndim)
    out = []
    cuts = [(0, ext[0] - 1)] if split is None else [(0, split - 1), (split, ext[0] - 1)]
    for lo, hi in cuts:
        b = P.Box()
        for a in range(3):
            b.lo[a], b.hi[a] = 0, ext[a] - 1
        b.lo[0], b.hi[0] = lo, hi
        out.append(b)
    return out


def _run(shape, split, periodic, sweeps, fused, seed=3):
    import torch
    from ampe_amd import hip as H
    ndim = len(shape)
    full = H.boxes_of(shape, ndim)
    lev = H.Level(ndim, _boxes(shape, split), 1, periodic=full if periodic else None)
    rng = np.random.default_rng(seed)
    dev = "cuda"
    fc = torch.from_numpy(-rng.uniform(0.5, 1.5, lev.nside_tot)).to(dev)
    sm = torch.from_numpy(lev.to_ghosted(rng.uniform(0.5, 1.5, shape))).to(dev)
    ncomp = 4
    rhs = torch.from_numpy(np.concatenate([lev.to_ghosted(rng.normal(size=shape))
                                           for _ in range(ncomp)])).to(dev)
    z = torch.from_numpy(np.concatenate([lev.to_ghosted(rng.normal(size=shape))
                                         for _ in range(ncomp)])).to(dev)
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    H.check(H.lib().ampe_ghost_fill(lev.h, ptr(sm), 1, None))
    dx = (C.c_double * 3)(0.5, 0.25, 0.75)
    H.check(H.lib().ampe_quat_smooth(lev.h, ptr(fc), ptr(sm), 0.3, dx, ncomp, ptr(rhs), ptr(z),
                                     sweeps, fused, None))
    zz = z.cpu().numpy().reshape(ncomp, -1)
    return np.stack([_interior(lev, zz[m], ndim) for m in range(ncomp)])


def _interior(lev, a, ndim):
    parts = []
    for p, b in enumerate(lev.boxes):
        n = [b.hi[x] - b.lo[x] + 1 for x in range(3)]
        if ndim == 2:
            n[2] = 1
        gz = 1 if ndim == 3 else 0
        shape = (n[2] + 2 * gz, n[1] + 2, n[0] + 2)
        blk = a[lev.cell_off[p]:lev.cell_off[p] + int(np.prod(shape))].reshape(shape)
        parts.append(blk[gz:shape[0] - gz, 1:-1, 1:-1].ravel())
    return np.concatenate(parts)


@pytest.mark.parametrize("shape,split,periodic", [
    ((96, 160), None, True),
    ((96, 160), 70, True),
    ((40, 48), 21, False),
    ((20, 36, 130), None, True),
    ((20, 36, 130), 66, True),
    ((12, 18, 70), 30, False),
])
@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_fused_sweep_bit_identical(dev, shape, split, periodic, sweeps):
    a = _run(shape, split, periodic, sweeps, fused=0)
    b = _run(shape, split, periodic, sweeps, fused=1)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)
