"""GPU parity of the regrid against the oracle (oracle/oracle_amr.c
o_hier_transfer, oracle/oracle.py regrid_levels):

- the data of a new hierarchy (ampe_hier_transfer: copies from the old same
  level, linear refinement of the new coarser level elsewhere) bit for bit,
  2-D and 3-D, periodic and bounded;
- new level boxes from the gradient detector and the clustering
  (ampe_hier_regrid) identical to the oracle's, up to three levels;
- a reduced dendrite run (anisotropic flux k = 4, nu = 0.05, dilute KKS with
  k_e = 0.14, bounded domain, uneven level-0 patches): the regrid every two
  steps on the device and in the oracle from their own states, tags of every
  level bit-exact from the same state, boxes identical, the integrator
  restarted on the new hierarchy at the current time, and the fields within
  1e-10 of the oracle's after every step.
"""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _p(t):
    return C.c_void_p(t.data_ptr())


def _interior(H):
    """Mask of the interior entries of a composite cell array."""
    m = np.zeros(H.ncg, dtype=np.uint8)
    for pp in range(H.np):
        H.ghost_view(m, pp)[(slice(1, -1),) * H.ndim] = 1
    return m.astype(bool)


TRANSFERS = {
    "2d_periodic": (2, (32, 32), True,
                    [[((0, 0), (15, 31)), ((16, 0), (31, 31))],
                     [((16, 16), (31, 39)), ((32, 20), (47, 35))]],
                    [[((0, 0), (15, 31)), ((16, 0), (31, 31))],
                     [((12, 14), (27, 29)), ((40, 40), (51, 55)), ((0, 0), (7, 7))],
                     [((30, 32), (45, 51))]]),
    "2d_bounded": (2, (32, 32), False,
                   [[((0, 0), (31, 31))], [((0, 0), (15, 11)), ((16, 16), (31, 39))]],
                   [[((0, 0), (31, 31))], [((0, 0), (23, 15)), ((24, 24), (45, 37))],
                    [((4, 4), (19, 15))]]),
    "3d_periodic": (3, (16, 16, 16), True,
                    [[((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 15, 15))],
                     [((8, 8, 8), (19, 15, 15))]],
                    [[((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 15, 15))],
                     [((4, 8, 6), (17, 17, 15)), ((20, 20, 20), (27, 25, 31))]]),
    "3d_bounded": (3, (16, 16, 16), False,
                   [[((0, 0, 0), (15, 15, 15))], [((0, 0, 0), (7, 7, 7))]],
                   [[((0, 0, 0), (15, 15, 15))], [((0, 0, 0), (11, 9, 7)), ((16, 16, 16), (23, 31, 27))],
                    [((2, 2, 2), (13, 9, 9))]]),
}


@pytest.mark.parametrize("case", list(TRANSFERS))
def test_transfer_matches_oracle(dev, case):
    import torch
    from ampe_amd import hip as H
    from oracle import oracle as O
    ndim, dom, periodic, old, new = TRANSFERS[case]
    print(f"transfer {case}", flush=True)
    ho, hn = O.Hier(ndim, old, dom, periodic=periodic), O.Hier(ndim, new, dom, periodic=periodic)
    go, gn = H.Hier(ndim, old, dom, periodic=periodic), H.Hier(ndim, new, dom, periodic=periodic)
    ncomp = 2
    f = _rng(3).uniform(-1.0, 1.0, ncomp * ho.ncg)
    want = O.hier_transfer(ho, hn, f, ncomp)
    fo = torch.from_numpy(f).to(dev)
    fn = torch.full((ncomp * gn.ncg,), float("nan"), dtype=torch.float64, device=dev)
    go.transfer_to(gn, _p(fo), _p(fn), ncomp)
    got = fn.cpu().numpy()
    im = np.tile(_interior(hn), ncomp)
    assert np.array_equal(got[im], want[im])


def _disk(n, ndim, r, w, seed):
    x = np.arange(n) + 0.5
    g = np.meshgrid(*([x] * ndim), indexing="ij")
    d = np.sqrt(sum((a - n / 2) ** 2 for a in g))
    phi = 0.5 * (1.0 - np.tanh((d - r) / (w / 2.0)))
    return np.clip(phi + 0.01 * _rng(seed).standard_normal(phi.shape) * (phi > 0.05) * (phi < 0.95),
                   0, 1)


def _cluster(eff=0.7, mn=2, mx=8, buf=1):
    cp = P.ClusterParams()
    cp.efficiency_tol = eff
    cp.min_size[:] = [mn] * 3
    cp.max_size[:] = [mx] * 3
    cp.tag_buffer = buf
    return cp


REGRIDS = {
    "2d_periodic": (2, 32, True, [[((0, 0), (19, 31)), ((20, 0), (31, 31))]]),
    "2d_bounded": (2, 32, False, [[((0, 0), (31, 31))], [((20, 20), (43, 35))]]),
    "3d_bounded": (3, 16, False, [[((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 15, 15))]]),
}


@pytest.mark.parametrize("case", list(REGRIDS))
def test_regrid_boxes_match_oracle(dev, case):
    import torch
    from ampe_amd import hip as H
    from oracle import oracle as O
    ndim, n, periodic, old = REGRIDS[case]
    print(f"regrid {case}", flush=True)
    dom = (n,) * ndim
    ho = O.Hier(ndim, old, dom, periodic=periodic)
    go = H.Hier(ndim, old, dom, periodic=periodic)
    f = ho.cells()
    phi = _disk(n, ndim, n / 5, 3.0, 11)
    cur = phi
    for lev in range(len(old)):
        ho.set_level(f, lev, cur)
        cur = np.kron(cur, np.ones((2,) * ndim))
    p = P.two_grain_params() if ndim == 2 else P.polycrystal_params(ndim=3)
    tol = 0.05
    cp = _cluster()
    want = O.regrid_levels(ho, f, p, tol, cp.efficiency_tol, tuple(cp.min_size),
                           tuple(cp.max_size), cp.tag_buffer, 3)
    assert len(want) == 3
    t = torch.from_numpy(f).to(dev)
    gn = go.regrid(_p(t), p, tol, cp, 3)
    assert gn.levels == [[(tuple(lo), tuple(hi)) for lo, hi in lev] for lev in want]


def _tags_equal(O, H, ho, hg, y, p0, tol, dev):
    """Tags of every level but the finest, device (ampe_tag_cells on the
    level) against the oracle, from the same composite phase field."""
    import torch
    f = np.ascontiguousarray(y[:ho.ncg])
    O.lib().o_hier_ghost_fill(ho.h, O.ptr(f), 1, 2)
    t = torch.from_numpy(f).to(dev)
    for lev in range(len(ho.levels)):
        want = O.hier_level_tags(ho, f, lev, p0 if lev == 0 else _lev_params(p0, lev), tol)
        tags = torch.zeros(hg.ncell, dtype=torch.int32, device=dev)
        H.check(H.lib().ampe_tag_cells(hg.level(lev), C.byref(_lev_params(p0, lev)), _p(t), tol,
                                       _p(tags), None), "ampe_tag_cells")
        got = tags.cpu().numpy()
        # device tags: interior cells per patch in the composite order
        off = 0
        for pp in range(ho.np):
            G = ho.geoms[pp]
            cnt = G.nx * G.ny * G.nz
            if ho.patch_level[pp] == lev:
                lo, _ = ho.boxes[pp]
                z0 = lo[2] if ho.ndim == 3 else 0
                blk = got[off:off + cnt].reshape(G.nz, G.ny, G.nx)
                w3 = want if ho.ndim == 3 else want[None]
                ref = w3[z0:z0 + G.nz, lo[1]:lo[1] + G.ny, lo[0]:lo[0] + G.nx]
                assert np.array_equal(blk, ref), (lev, pp)
            off += cnt


def _lev_params(p0, lev):
    p = type(p0)()
    C.memmove(C.byref(p), C.byref(p0), C.sizeof(p0))
    for a in range(3):
        p.dx[a] = p0.dx[a] / (1 << lev)
    return p


def test_dendrite_regrid_run_matches_oracle(dev):
    """Reduced config-4 run: 32^2 level 0 in two uneven patches, up to two
    levels, the fine level built by the gradient detector at t = 0 and
    rebuilt every two steps; device and oracle regrid from their own states."""
    import torch
    from ampe_amd import hip as H
    from oracle import oracle as O
    n, ndim, periodic = 32, 2, False
    lev0 = [((0, 0), (19, 31)), ((20, 0), (31, 31))]
    p = P.dilute_params(ndim=2, dilute_keq=0.14)
    p.flux_type = P.FLUX_ANISOTROPIC
    p.aniso_nu = 0.05
    p.aniso_knumber = 4
    s = P.solver_defaults(dt=1e-7, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8)
    phi = _disk(n, 2, 6.0, 3.0, 7)
    c = 0.05 + 0.002 * _rng(8).uniform(-1, 1, (n, n))
    q = np.zeros((4, n, n))
    q[0] = 1.0
    fields = [phi, c] + list(q)
    nc = len(fields)
    # cluster parameters change between regrids so that the fine patches move
    # (new cells refined from the coarse level next to copied old ones)
    tol, max_levels = 0.05, 2
    cps = [_cluster(0.7, 2, 12, 1), _cluster(0.8, 2, 8, 2), _cluster(0.6, 2, 16, 1)]
    dx0 = tuple(p.dx)
    ho = O.Hier(ndim, [lev0], (n, n), dx0, periodic)
    hg = H.Hier(ndim, [lev0], (n, n), dx0, periodic)
    yo = ho.cells(nc)
    for k, f in enumerate(fields):
        ho.set_level(yo, 0, f, comp=k)
    yg = yo.copy()
    t_now, so, sg = 0.0, None, None
    for cycle in range(3):
        print(f"dendrite regrid cycle {cycle} t {t_now:.3e}", flush=True)
        cp = cps[cycle]
        if cycle > 0:
            _tags_equal(O, H, ho, hg, yg, p, tol, dev)
        # regrid both from their own states
        lv = O.regrid_levels(ho, yo[:ho.ncg], p, tol, cp.efficiency_tol, tuple(cp.min_size),
                             tuple(cp.max_size), cp.tag_buffer, max_levels)
        hon = O.Hier(ndim, lv, (n, n), dx0, periodic)
        yon = O.hier_transfer(ho, hon, yo, nc)
        tg = torch.from_numpy(np.ascontiguousarray(yg)).to(dev)
        hgn = hg.regrid(_p(tg), p, tol, cp, max_levels)
        assert hgn.levels == hon.levels, (cycle, hgn.levels, hon.levels)
        assert len(hgn.levels) == 2
        tn = torch.zeros(nc * hgn.ncg, dtype=torch.float64, device=dev)
        hg.transfer_to(hgn, _p(tg), _p(tn), nc)
        ygn = tn.cpu().numpy()
        ho, hg, yo, yg = hon, hgn, yon, ygn
        so, sg = O.HierSim(p, s, ho), H.HierSim(p, s, hg)
        so.set_state(yo)
        sg.set_state(yg)
        so_t = t_now
        if cycle > 0:
            so.set_time(t_now)
            sg.set_time(t_now)
        im = _interior(ho)
        for step in range(2):
            print(f"  step {step}", flush=True)
            ro, sto = so.advance(1)
            rg, stg = sg.advance(1)
            assert ro == rg == 0, (cycle, step, sto, stg)
            for k in ("nst", "nni", "nli", "ncfn", "nlcf"):
                assert sto[k] == stg[k], (cycle, step, k, sto, stg)
            assert abs(sto["t"] - stg["t"]) <= 1e-12 * sto["t"]
            yo, yg = so.get_state(), sg.get_state()
            for m in range(nc):
                a = yg[m * ho.ncg:(m + 1) * ho.ncg][im]
                b = yo[m * ho.ncg:(m + 1) * ho.ncg][im]
                assert _rel(a, b) < 1e-10, (cycle, step, m, _rel(a, b))
            t_now = sto["t"]
        assert t_now > so_t
