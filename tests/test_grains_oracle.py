"""CPU checks of the grain numbering restated in the oracle
(Grains::findAndNumberGrains): closed cases (separate blobs, periodic wrap,
corner contact, a U-shaped grain that needs several sweeps) and random fields
against an independent labelling (scipy.ndimage.label with face connectivity,
components merged across the periodic boundaries, numbered by their smallest
x-fastest index)."""
import ctypes as C

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O


def _label(phase, thr=0.5, periodic=True):
    """phase: (nz, ny, nx) or (ny, nx); returns (labels, ngrains, sweeps)."""
    ndim = phase.ndim
    shape = phase.shape if ndim == 3 else (1,) + phase.shape
    pt = O.Patch(ndim, (shape[2], shape[1], shape[0]))
    a = pt.load(pt.cells(), phase.reshape(shape))
    lab = np.zeros(pt.N, dtype=np.int32)
    sw = C.c_int(0)
    ng = O.lib().o_grains_label(C.byref(pt.G), O.ptr(a), thr, int(periodic), O.iptr(lab),
                                C.byref(sw))
    return lab.reshape(shape if ndim == 3 else shape[1:]), ng, sw.value


def _reference_labels(mask):
    """Independent periodic face-connected labelling, numbered by each
    component's smallest flat (x-fastest) index."""
    ndim = mask.ndim
    lab, n = ndimage.label(mask, structure=ndimage.generate_binary_structure(ndim, 1))
    parent = list(range(n + 1))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for ax in range(ndim):
        a = np.take(lab, 0, axis=ax)
        b = np.take(lab, lab.shape[ax] - 1, axis=ax)
        for u, v in zip(a.ravel(), b.ravel()):
            if u and v:
                parent[find(u)] = find(v)
    root = np.array([find(x) for x in range(n + 1)])
    comp = root[lab]
    flat = comp.ravel()
    first = {}
    for idx in np.flatnonzero(flat):
        r = flat[idx]
        if r not in first:
            first[r] = idx
    order = sorted(first, key=lambda r: first[r])
    rank = {r: k for k, r in enumerate(order)}
    out = np.full(flat.shape, -1, dtype=np.int32)
    nz = np.flatnonzero(flat)
    out[nz] = [rank[r] for r in flat[nz]]
    return out.reshape(mask.shape), len(order)


def test_two_blobs_and_order():
    f = np.zeros((16, 20))
    f[10:13, 2:5] = 1.0   # smallest flat index 10*20+2 = 202
    f[2:5, 12:16] = 1.0   # smallest flat index 2*20+12 = 52 -> grain 0
    lab, ng, _ = _label(f)
    assert ng == 2
    assert (lab[2:5, 12:16] == 0).all() and (lab[10:13, 2:5] == 1).all()
    assert (lab[f < 0.5] == -1).all()


def test_periodic_wrap_joins():
    f = np.zeros((12, 12))
    f[4:7, 0:2] = 1.0
    f[4:7, 10:12] = 1.0
    lab, ng, _ = _label(f)
    assert ng == 1
    _, ng_np, _ = _label(f, periodic=False)  # without the wrap: two grains
    assert ng_np == 2


def test_corner_contact_is_not_connected():
    f = np.zeros((10, 10))
    f[2, 2] = f[3, 3] = 1.0
    _, ng, _ = _label(f)
    assert ng == 2
    f3 = np.zeros((6, 6, 6))
    f3[1, 1, 1] = f3[2, 2, 1] = 1.0  # edge contact in 3-D
    _, ng3, _ = _label(f3)
    assert ng3 == 2


def test_u_shape_needs_several_sweeps():
    f = np.zeros((20, 20))
    f[2:18, 2] = f[2:18, 17] = 1.0
    f[17, 2:18] = 1.0  # U open at the top: the smallest index is at (2,2)
    lab, ng, sweeps = _label(f)
    assert ng == 1 and sweeps > 2
    assert (lab[f >= 0.5] == 0).all()


@pytest.mark.parametrize("shape,seed,thr", [((40, 36), 1, 0.55), ((18, 20, 22), 2, 0.7)])
def test_random_fields_match_independent_labelling(shape, seed, thr):
    rng = np.random.default_rng(seed)
    f = ndimage.gaussian_filter(rng.normal(size=shape), 1.5, mode="wrap")
    f = (f - f.min()) / (f.max() - f.min())
    lab, ng, _ = _label(f, thr)
    ref, nref = _reference_labels(f >= thr)
    assert ng == nref and ng > 3
    assert np.array_equal(lab, ref)
