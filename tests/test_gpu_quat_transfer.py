"""Orientation-field transfers on the device (ampe_quat_coarsen /
ampe_quat_refine) against the oracle's restatement of quatcoarsen /
quatlinrefine (oracle/quat_transfer.py) on periodic one- and two-patch levels,
plus closed forms: a uniform unit quaternion is unchanged, every result has
unit norm."""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu

OQ = pytest.importorskip("oracle.quat_transfer")


def _levels(shape, split):
    from ampe_amd import hip as H
    ndim = len(shape)
    full = H.boxes_of(shape, ndim)
    cshape = tuple(s // 2 for s in shape)
    cfull = H.boxes_of(cshape, ndim)
    if not split:
        return H.Level(ndim, [full], 1, periodic=full), H.Level(ndim, [cfull], 1, periodic=cfull)
    cut = shape[-1] // 2 + 2  # even, uneven halves
    fb, cb = [], []
    for lo, hi in ((0, cut - 1), (cut, shape[-1] - 1)):
        b, c = H.Box(), H.Box()
        for d in range(3):
            b.lo[d], b.hi[d] = full.lo[d], full.hi[d]
            c.lo[d], c.hi[d] = cfull.lo[d], cfull.hi[d]
        b.lo[0], b.hi[0] = lo, hi
        c.lo[0], c.hi[0] = lo // 2, (hi + 1) // 2 - 1
        fb.append(b)
        cb.append(c)
    return H.Level(ndim, fb, 1, periodic=full), H.Level(ndim, cb, 1, periodic=cfull)


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    idx = np.indices(shape).astype(float)
    q = np.stack([np.cos(0.3 * idx.sum(0) + rng.uniform(0, 6)) + 0.1 * rng.normal(size=shape)
                  for _ in range(4)])
    return q / np.sqrt((q ** 2).sum(0))


def _dev(lev, arr):
    import torch
    return torch.from_numpy(np.concatenate([lev.to_ghosted(a) for a in arr])).to("cuda")


def _host(lev, t, ql, shape):
    v = t.cpu().numpy().reshape(ql, -1)
    out = np.zeros((ql,) + shape)
    for m in range(ql):
        for p in range(len(lev.boxes)):
            blk = lev.ghost_view(v[m], p)
            b = lev.boxes[p]
            nd = len(shape)
            g = 1
            gz = g if nd == 3 else 0
            n = lev._patch_ext(b)
            sl = tuple(slice(b.lo[d], b.lo[d] + n[d]) for d in reversed(range(nd)))
            inner = blk[gz:gz + n[2], g:g + n[1], g:g + n[0]] if nd == 3 else blk[g:g + n[1], g:g + n[0]]
            out[(m,) + sl] = inner
    return out


@pytest.mark.parametrize("shape", [(32, 48), (16, 24, 20)])
@pytest.mark.parametrize("split", [False, True])
def test_quat_coarsen_refine_match_oracle(dev, shape, split):
    from ampe_amd import hip as H
    ndim = len(shape)
    dx = (C.c_double * 3)(0.05, 0.05, 0.05)
    fine, coarse = _levels(shape, split)
    cshape = tuple(s // 2 for s in shape)
    qf = _field(shape, 3)
    tf = _dev(fine, qf)
    import torch
    tc = torch.zeros(4 * coarse.ncg, dtype=torch.float64, device="cuda")
    H.check(H.lib().ampe_quat_coarsen(coarse.h, fine.h, dx, 4, C.c_void_p(tf.data_ptr()),
                                      C.c_void_p(tc.data_ptr()), None))
    got_c = _host(coarse, tc, 4, cshape)
    ref_c = OQ.coarsen(qf, (0.05, 0.05, 0.05))
    assert np.max(np.abs(got_c - ref_c)) <= 1e-15
    assert np.allclose((got_c ** 2).sum(0), 1.0, atol=1e-14)
    # refine the oracle's coarse field (ghosts filled first)
    qc = OQ.coarsen(_field(shape, 4), (0.05, 0.05, 0.05))
    tc = _dev(coarse, qc)
    H.check(H.lib().ampe_ghost_fill(coarse.h, C.c_void_p(tc.data_ptr()), 4, None))
    tf2 = torch.zeros(4 * fine.ncg, dtype=torch.float64, device="cuda")
    H.check(H.lib().ampe_quat_refine(fine.h, coarse.h, dx, 4, C.c_void_p(tc.data_ptr()),
                                     C.c_void_p(tf2.data_ptr()), None))
    got_f = _host(fine, tf2, 4, shape)
    ref_f = OQ.refine(qc, (0.05, 0.05, 0.05))
    assert np.max(np.abs(got_f - ref_f)) <= 1e-15
    assert np.allclose((got_f ** 2).sum(0), 1.0, atol=1e-14)


def test_uniform_quaternion_unchanged(dev):
    from ampe_amd import hip as H
    import torch
    shape = (16, 16)
    fine, coarse = _levels(shape, False)
    q0 = np.array([0.5, -0.5, 0.5, 0.5])
    qf = np.broadcast_to(q0[:, None, None], (4,) + shape).copy()
    dx = (C.c_double * 3)(1.0, 1.0, 1.0)
    tf = _dev(fine, qf)
    tc = torch.zeros(4 * coarse.ncg, dtype=torch.float64, device="cuda")
    H.check(H.lib().ampe_quat_coarsen(coarse.h, fine.h, dx, 4, C.c_void_p(tf.data_ptr()),
                                      C.c_void_p(tc.data_ptr()), None))
    H.check(H.lib().ampe_ghost_fill(coarse.h, C.c_void_p(tc.data_ptr()), 4, None))
    tf2 = torch.zeros_like(tf)
    H.check(H.lib().ampe_quat_refine(fine.h, coarse.h, dx, 4, C.c_void_p(tc.data_ptr()),
                                     C.c_void_p(tf2.data_ptr()), None))
    got = _host(fine, tf2, 4, shape)
    assert np.max(np.abs(got - qf)) < 1e-15


@pytest.mark.parametrize("shape", [(32, 48), (16, 24, 20)])
@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("ncomp", [1, 6])
def test_linear_refine_matches_oracle(dev, shape, split, ncomp):
    """ampe_linear_refine (scalar fields: the interpolation without the
    normalisation) against the oracle, any number of components."""
    import torch
    from ampe_amd import hip as H
    print("case", shape, split, ncomp, flush=True)
    fine, coarse = _levels(shape, split)
    cshape = tuple(s // 2 for s in shape)
    rng = np.random.default_rng(7 + ncomp)
    vc = rng.uniform(-1.0, 2.0, (ncomp,) + cshape)
    tc = _dev(coarse, vc)
    H.check(H.lib().ampe_ghost_fill(coarse.h, C.c_void_p(tc.data_ptr()), ncomp, None))
    tf = torch.zeros(ncomp * fine.ncg, dtype=torch.float64, device="cuda")
    dx = (C.c_double * 3)(0.05, 0.1, 0.2)
    H.check(H.lib().ampe_linear_refine(fine.h, coarse.h, dx, ncomp, C.c_void_p(tc.data_ptr()),
                                       C.c_void_p(tf.data_ptr()), None))
    got = _host(fine, tf, ncomp, shape)
    ref = OQ.refine(vc, (0.05, 0.1, 0.2), normalize=False)
    assert np.max(np.abs(got - ref)) <= 1e-15
