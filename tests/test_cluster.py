"""Tag clustering (ampe_cluster_mask, host code of the library) against the
oracle's restatement of Berger-Rigoutsos (oracle/cluster.py), closed-form
cases and the invariants of a box generator: buffered tags covered, boxes
disjoint, inside the domain, no larger than max_size, and every box either
efficient enough or not cuttable."""
import itertools

import numpy as np
import pytest
from scipy import ndimage

from ampe_amd import params as P

cl = pytest.importorskip("oracle.cluster")


def _cp(eff=0.8, mn=2, mx=0, buf=0):
    cp = P.ClusterParams()
    cp.efficiency_tol = eff
    for d in range(3):
        cp.min_size[d] = mn
        cp.max_size[d] = mx
    cp.tag_buffer = buf
    return cp


def _lib(mask, cp, lo=None):
    from ampe_amd import hip as H
    lo = lo or (0, 0, 0)
    nd = mask.ndim
    return [(b[0][:nd], b[1][:nd]) for b in H.cluster_mask(mask, cp, lo=lo)]


def _buffered(mask, b):
    if b == 0:
        return mask != 0
    return ndimage.binary_dilation(mask != 0, structure=np.ones((3,) * mask.ndim),
                                   iterations=b)


def _check_invariants(mask, boxes, eff, mx, buf, mn):
    nd = mask.ndim
    tags = _buffered(mask, buf)
    cover = np.zeros(mask.shape, np.int32)
    for lo, hi in boxes:
        sl = tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(nd)))
        assert all(0 <= lo[d] <= hi[d] < mask.shape[nd - 1 - d] for d in range(nd))
        if mx > 0:
            assert all(hi[d] - lo[d] + 1 <= mx for d in range(nd))
        cover[sl] += 1
    assert cover.max() <= 1, "boxes overlap"
    assert np.all(cover[tags] == 1), "tag outside every box"
    if mx == 0:
        for lo, hi in boxes:
            sl = tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(nd)))
            n = tags[sl].sum()
            ext = [hi[d] - lo[d] + 1 for d in range(nd)]
            assert n >= eff * np.prod(ext) or all(e < 2 * mn for e in ext)


def test_closed_forms():
    m = np.zeros((32, 40), np.uint8)
    assert _lib(m, _cp()) == []
    m[5:9, 3:15] = 1
    m[20:30, 25:31] = 1
    boxes = sorted(_lib(m, _cp(eff=1.0)))
    assert boxes == [((3, 5), (14, 8)), ((25, 20), (30, 29))]
    full = np.ones((8, 8, 8), np.uint8)
    assert _lib(full, _cp()) == [((0, 0, 0), (7, 7, 7))]
    one = np.zeros((16, 16), np.uint8)
    one[7, 9] = 1
    assert _lib(one, _cp(), lo=(100, 200, 0)) == [((109, 207), (109, 207))]
    # an L shape at efficiency 1 is cut into exactly two full rectangles
    L = np.zeros((20, 20), np.uint8)
    L[2:18, 2:6] = 1
    L[14:18, 6:16] = 1
    bl = _lib(L, _cp(eff=1.0, mn=1))
    assert len(bl) == 2 and sum((h[0] - l[0] + 1) * (h[1] - l[1] + 1) for l, h in bl) == L.sum()


@pytest.mark.parametrize("ndim,seed,eff,mn,mx,buf", list(itertools.product(
    [2, 3], [1, 2], [0.6, 0.9], [1, 3], [0, 12], [0, 1])))
def test_matches_oracle_and_invariants(ndim, seed, eff, mn, mx, buf):
    rng = np.random.default_rng(seed)
    shape = (48, 56) if ndim == 2 else (20, 24, 28)
    f = ndimage.gaussian_filter(rng.normal(size=shape), 2.5)
    mask = (f > np.quantile(f, 0.85)).astype(np.uint8)
    got = _lib(mask, _cp(eff, mn, mx, buf))
    ref = cl.cluster(mask, eff, (mn,) * 3, (mx,) * 3, buf)
    assert got == ref
    _check_invariants(mask, got, eff, mx, buf, mn)
