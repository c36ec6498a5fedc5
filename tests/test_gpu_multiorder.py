"""Multi-order phase right-hand side and projection on the device against
the oracle (MultiOrderRHSStrategy::evaluateRHS, computerhsmultiorder,
projectphi), on a periodic level of two patches and on a single bounded
patch, with the uniform mobility and with a mobility field."""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu

GEOMS = {2: (40, 24, 1), 3: (20, 12, 10)}


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("layout", ["periodic_two_patches", "bounded"])
@pytest.mark.parametrize("norder", [2, 5])
def test_multiorder_rhs_and_projection(dev, ndim, layout, norder):
    import torch
    from oracle import oracle as O
    from ampe_amd import hip as H
    p = P.two_grain_params(ndim=ndim, dx=(0.5, 0.25, 0.125))
    nx, ny, nz = GEOMS[ndim]
    print("case", ndim, layout, norder, flush=True)
    pt = O.Patch(ndim, (nx, ny, nz), dx=tuple(p.dx))
    pt.G.periodic = 1 if layout.startswith("periodic") else 0
    rng = np.random.Generator(np.random.Philox(41 + norder))
    shape = (pt.nz, pt.ny, pt.nx)
    vals = rng.uniform(0.0, 1.0, (norder,) + shape)
    mob = rng.uniform(0.5, 2.0, shape)
    phi, M = pt.cells(norder), pt.cells()
    pt.interior(phi)[...] = vals
    pt.interior(M)[...] = mob
    O.lib().o_ghost_fill(C.byref(pt.G), O.ptr(phi), norder)
    hi = [nx - 1, ny - 1, nz - 1 if ndim == 3 else 0]
    if layout == "periodic_two_patches":
        b0, b1 = P.Box(), P.Box()
        b0.lo[:] = [0, 0, 0]
        b0.hi[:] = [hi[0], ny // 2 - 1, hi[2]]
        b1.lo[:] = [0, ny // 2, 0]
        b1.hi[:] = hi
        boxes = [b0, b1]
        dom = P.Box()
        dom.lo[:] = [0, 0, 0]
        dom.hi[:] = hi
        lev = H.Level(ndim, boxes, 1, periodic=dom)
    else:
        b = P.Box()
        b.lo[:] = [0, 0, 0]
        b.hi[:] = hi
        boxes = [b]
        lev = H.Level(ndim, boxes, 1, periodic=None)
    sq = (lambda a: a) if ndim == 3 else (lambda a: a[0])
    ph = np.concatenate([lev.to_ghosted(sq(vals[i])) for i in range(norder)])
    phd = torch.from_numpy(ph).to(dev)
    H.check(H.lib().ampe_ghost_fill(lev.h, C.c_void_p(phd.data_ptr()), norder, None))
    Md = torch.from_numpy(lev.to_ghosted(sq(mob))).to(dev)
    rd = torch.zeros(norder * lev.ncg, dtype=torch.float64, device=dev)

    def domain(buf, comp):
        out = np.zeros(shape)
        for q, bx in enumerate(boxes):
            blk = lev.ghost_view(buf[comp * lev.ncg:(comp + 1) * lev.ncg], q)
            blk = blk if ndim == 3 else blk[None]
            n = [bx.hi[d] - bx.lo[d] + 1 for d in range(3)]
            gz = 1 if ndim == 3 else 0
            out[bx.lo[2]:bx.lo[2] + n[2], bx.lo[1]:bx.lo[1] + n[1], :] = \
                blk[gz:gz + n[2], 1:1 + n[1], 1:1 + n[0]]
        return out

    for Mo, Mp in ((None, None), (M, Md)):
        r_o = pt.cells(norder)
        O.lib().o_phase_rhs_multiorder(C.byref(pt.G), C.byref(p), norder, O.ptr(phi),
                                       O.ptr(Mo) if Mo is not None else None, O.ptr(r_o))
        H.check(H.lib().ampe_phase_rhs_multiorder(
            lev.h, C.byref(p), norder, C.c_void_p(phd.data_ptr()),
            C.c_void_p(Mp.data_ptr()) if Mp is not None else None,
            C.c_void_p(rd.data_ptr()), None))
        rh = rd.cpu().numpy()
        for i in range(norder):
            assert _rel(domain(rh, i), pt.interior(r_o[i])) < 1e-13, (i, Mo is None)
    # projection
    err_v = rng.normal(size=(norder,) + shape)
    err_o, corr_o = pt.cells(norder), pt.cells(norder)
    pt.interior(err_o)[...] = err_v
    O.lib().o_project_phases(C.byref(pt.G), norder, O.ptr(phi), O.ptr(corr_o), O.ptr(err_o))
    ed = torch.from_numpy(np.concatenate([lev.to_ghosted(sq(err_v[i]))
                                          for i in range(norder)])).to(dev)
    cd = torch.zeros_like(ed)
    H.check(H.lib().ampe_project_phases(lev.h, norder, C.c_void_p(phd.data_ptr()),
                                        C.c_void_p(cd.data_ptr()), C.c_void_p(ed.data_ptr()),
                                        None))
    ch, eh = cd.cpu().numpy(), ed.cpu().numpy()
    for i in range(norder):
        assert np.array_equal(domain(ch, i), pt.interior(corr_o[i]))
        assert np.array_equal(domain(eh, i), pt.interior(err_o[i]))


def test_multiorder_rejects_bad_arguments(dev):
    from ampe_amd import hip as H
    p = P.two_grain_params(ndim=2)
    b = P.Box()
    b.lo[:] = [0, 0, 0]
    b.hi[:] = [7, 7, 0]
    lev = H.Level(2, [b], 1, periodic=b)
    import torch
    t = torch.zeros(4 * lev.ncg, dtype=torch.float64, device=dev)
    ptr = C.c_void_p(t.data_ptr())
    assert H.lib().ampe_phase_rhs_multiorder(lev.h, C.byref(p), 1, ptr, None, ptr, None) != 0
    p.flux_type = P.FLUX_ISOTROPIC
    assert H.lib().ampe_phase_rhs_multiorder(lev.h, C.byref(p), 2, ptr, None, ptr, None) != 0
