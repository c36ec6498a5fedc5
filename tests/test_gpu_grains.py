"""GPU grain numbering (ampe_grains_label) against the oracle's restatement
of Grains::findAndNumberGrains: bit-exact labels on smooth random fields
(2-D and 3-D, periodic), on the polycrystal state, and the same labels when
the level is split into uneven patches."""
import ctypes as C

import numpy as np
import pytest
from scipy import ndimage

from ampe_amd import params as P

pytestmark = pytest.mark.gpu


def _oracle(phase, thr, periodic=True):
    from oracle import oracle as O
    ndim = phase.ndim
    shape = phase.shape if ndim == 3 else (1,) + phase.shape
    pt = O.Patch(ndim, (shape[2], shape[1], shape[0]))
    a = pt.load(pt.cells(), phase.reshape(shape))
    lab = np.zeros(pt.N, dtype=np.int32)
    sw = C.c_int(0)
    ng = O.lib().o_grains_label(C.byref(pt.G), O.ptr(a), thr, int(periodic), O.iptr(lab),
                                C.byref(sw))
    return lab.reshape(phase.shape), ng


def _gpu(phase, thr, boxes=None, periodic=True):
    import torch
    from ampe_amd import hip as H
    ndim = phase.ndim
    full = H.boxes_of(phase.shape, ndim)
    lev = H.Level(ndim, boxes or [full], 1, periodic=full if periodic else None)
    f = torch.from_numpy(lev.to_ghosted(phase)).to("cuda")
    lab = torch.full((lev.ncell,), -7, dtype=torch.int32, device="cuda")
    ng = C.c_int(-1)
    H.check(H.lib().ampe_grains_label(lev.h, C.c_void_p(f.data_ptr()), thr,
                                      C.c_void_p(lab.data_ptr()), C.byref(ng), None))
    return lev.from_interior(lab.cpu().numpy()), ng.value


def _field(shape, seed, sigma=1.5):
    rng = np.random.default_rng(seed)
    f = ndimage.gaussian_filter(rng.normal(size=shape), sigma, mode="wrap")
    return (f - f.min()) / (f.max() - f.min())


@pytest.mark.parametrize("shape,seed,thr", [((128, 128), 3, 0.6), ((40, 44, 48), 4, 0.68)])
def test_labels_bit_exact(dev, shape, seed, thr):
    f = _field(shape, seed)
    lo, no = _oracle(f, thr)
    lg, ng = _gpu(f, thr)
    assert ng == no and no > 3
    assert np.array_equal(lg, lo)


@pytest.mark.parametrize("ndim", [2, 3])
def test_non_periodic_domain(dev, ndim):
    """Physical boundaries (no ghost fill from the other side), as in the
    reference's ComputeNumberOfGrains deck (periodic_dimension = 0, 0)."""
    shape = (30, 40) if ndim == 2 else (20, 24, 28)
    f = _field(shape, 21)
    thr = 0.6 if ndim == 2 else 0.66
    lo, no = _oracle(f, thr, periodic=False)
    lg, ng = _gpu(f, thr, periodic=False)
    lp, npp = _oracle(f, thr, periodic=True)
    assert ng == no and no > 2
    assert np.array_equal(lg, lo)
    assert npp <= no


def test_polycrystal_state(dev):
    phi, _, _ = P.polycrystal_state(n=48, seed=7, ngrains=8, solid_fraction=0.4)
    lo, no = _oracle(phi, 0.5)
    lg, ng = _gpu(phi, 0.5)
    assert ng == no and no >= 1
    assert np.array_equal(lg, lo)


@pytest.mark.parametrize("ndim", [2, 3])
def test_uneven_patches_same_labels(dev, ndim):
    from ampe_amd import hip as H
    shape = (64, 56) if ndim == 2 else (24, 28, 30)
    f = _field(shape, 11)
    thr = 0.6 if ndim == 2 else 0.68
    one, n1 = _gpu(f, thr)
    nx = shape[-1]
    bs = []
    for lo, hi in ((0, 19), (20, nx - 1)):
        b = P.Box()
        for a in range(3):
            b.lo[a] = 0
            b.hi[a] = (shape[::-1][a] - 1) if a < ndim else 0
        b.lo[0], b.hi[0] = lo, hi
        bs.append(b)
    two, n2 = _gpu(f, thr, bs)
    assert n1 == n2 and n1 > 2
    assert np.array_equal(one, two)


@pytest.mark.gpu
@pytest.mark.parametrize("ndim", [2, 3])
def test_cluster_tags_of_tagged_level(dev, ndim):
    """Device tags from the gradient detector (ampe_tag_cells) of a level of
    two patches, clustered by ampe_cluster_tags, give the boxes of the
    oracle's clustering of the same tags."""
    import torch
    from ampe_amd import hip as H, params as P
    from oracle import cluster as OC
    n = 64 if ndim == 2 else 24
    shape = (n,) * ndim
    idx = np.indices(shape).astype(float) + 0.5
    r = np.sqrt(sum((x - 0.4 * n) ** 2 for x in idx))
    phi = 0.5 * (1 - np.tanh((r - 0.25 * n) / 2.0))
    full = H.boxes_of(shape, ndim)
    half = n // 2
    b0, b1 = H.Box(), H.Box()
    for d in range(3):
        b0.lo[d], b0.hi[d] = full.lo[d], full.hi[d]
        b1.lo[d], b1.hi[d] = full.lo[d], full.hi[d]
    b0.hi[0] = half - 1
    b1.lo[0] = half
    lev = H.Level(ndim, [b0, b1], 1, periodic=full)
    p = P.two_grain_params(ndim=ndim)
    f = torch.from_numpy(lev.to_ghosted(phi)).to("cuda")
    H.check(H.lib().ampe_ghost_fill(lev.h, C.c_void_p(f.data_ptr()), 1, None))
    tags = torch.zeros(lev.ncell, dtype=torch.int32, device="cuda")
    H.check(H.lib().ampe_tag_cells(lev.h, C.byref(p), C.c_void_p(f.data_ptr()), 0.1,
                                   C.c_void_p(tags.data_ptr()), None))
    cp = P.ClusterParams()
    cp.efficiency_tol = 0.7
    for d in range(3):
        cp.min_size[d] = 2
        cp.max_size[d] = 16
    cp.tag_buffer = 1
    out = (H.Box * 4096)()
    nb = C.c_int(0)
    H.check(H.lib().ampe_cluster_tags(lev.h, C.c_void_p(tags.data_ptr()), C.byref(cp), out,
                                      4096, C.byref(nb), None))
    got = [(tuple(out[b].lo)[:ndim], tuple(out[b].hi)[:ndim]) for b in range(nb.value)]
    mask = lev.from_interior(tags.cpu().numpy())
    assert mask.sum() > 0
    ref = OC.cluster(mask, 0.7, (2,) * 3, (16,) * 3, 1)
    assert got == ref and len(got) > 1
