"""Pins of the oracle's hierarchy restatement (oracle/oracle_amr.c) that do
not retype its formulas:

- linear fields: linear refinement of the coarse-fine ghosts and the
  conservative coarsening are exact, constant refinement takes the parent;
- the control-volume weights add up to the level-0 cell count;
- composite conservation: with the fine fluxes coarsened onto the coarse-fine
  faces the composition right-hand side integrates to zero on a periodic
  hierarchy, and it does not when the levels are evaluated separately;
- a fine level covering the whole domain reduces the composite right-hand
  side to the single-level one on the fine grid (the single-patch oracle,
  another code path);
- Ewing's coarse-fine fix makes the composite operator exact on linear
  fields (its weights 2/(1+r), 1/(1+r) and the mirror cell; the reference's
  single-precision weights leave ~1e-8), while constant refinement alone
  leaves an O(1) residual next to the coarse-fine boundary;
- the FAC cycles over levels converge, and with the fine level covering the
  domain reach the fine single-level solution;
- the hierarchy integrator on one level is the single-patch integrator (same
  steps, fields to rounding), and with a fine level covering the domain it
  follows the single-patch integrator on the fine grid.

The FAC composite residual replaces only the coarse faces on fine patch
boundaries (EllipticFACOps.cc:1524-1528 coarsens OutersideData); its effect on
covered cells, which the middle levels of a three-level cycle restrict, is
pinned by the device parity tests, not here.
"""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P
from oracle import oracle as O

LEV0 = [((0, 0), (15, 31)), ((16, 0), (31, 31))]
LEV1 = [((16, 16), (31, 39)), ((32, 20), (47, 35))]


def _lin(n, h, a=0.3, b=0.7, c=-0.2):
    y, x = np.mgrid[0:n, 0:n]
    return a + b * (x + 0.5) * h + c * (y + 0.5) * h


def _fill_levels(H, fn):
    a = H.cells()
    for lev in range(len(H.levels)):
        n = H.dom0[0] << lev
        H.set_level(a, lev, fn(lev, n))
    return a


def test_weights_and_linear_transfers():
    H = O.Hier(2, [LEV0, LEV1], (32, 32), periodic=False)
    w = H.weights()
    assert w.sum() == 32 * 32
    a = _fill_levels(H, lambda l, n: _lin(n, 1.0 / (1 << l)))
    b = a.copy()
    O.lib().o_hier_ghost_fill(H.h, O.ptr(b), 1, 2)
    c = a.copy()
    O.lib().o_hier_ghost_fill(H.h, O.ptr(c), 1, 1)
    for p in range(H.np):
        if H.patch_level[p] != 1:
            continue
        (lo, hi) = H.boxes[p]
        gv, gc = H.ghost_view(b, p), H.ghost_view(c, p)
        for j in range(gv.shape[0]):
            for i in range(gv.shape[1]):
                gi, gj = lo[0] + i - 1, lo[1] + j - 1
                if not (0 <= gi < 64 and 0 <= gj < 64):
                    continue
                fine_on_level = any(l0[0] <= gi <= h0[0] and l0[1] <= gj <= h0[1]
                                    for (l0, h0) in H.levels[1])
                exact = 0.3 + 0.7 * (gi + 0.5) * 0.5 - 0.2 * (gj + 0.5) * 0.5
                assert abs(gv[j, i] - exact) < 1e-14
                if not fine_on_level:
                    parent = 0.3 + 0.7 * (gi // 2 + 0.5) - 0.2 * (gj // 2 + 0.5)
                    assert abs(gc[j, i] - parent) < 1e-14
    d = a.copy()
    O.lib().o_hier_coarsen(H.h, O.ptr(d), 1, 0)
    assert np.max(np.abs(H.get_level(d, 0) - _lin(32, 1.0))) < 1e-14


def _conc_setup(H, seed):
    p = P.two_grain_params()
    rng = np.random.default_rng(seed)
    c = _fill_levels(H, lambda l, n: 0.4 + 0.05 * rng.standard_normal((n, n)))
    ph = _fill_levels(H, lambda l, n: rng.uniform(0, 1, (n, n)))
    for f in (c, ph):
        O.lib().o_hier_coarsen(H.h, O.ptr(f), 1, 0)
        O.lib().o_hier_ghost_fill(H.h, O.ptr(f), 1, 2)
    cl, cs = c.copy(), 1.1 * c
    return p, c, ph, cl, cs


def test_composite_conservation_needs_flux_coarsening():
    H = O.Hier(2, [LEV0, LEV1], (32, 32), periodic=True)
    p, c, ph, cl, cs = _conc_setup(H, 1)
    Dc, Dp = H.sides(), H.sides()
    O.lib().o_hier_conc_coeffs(H.h, C.byref(p), O.ptr(ph), O.ptr(cl), O.ptr(cs), O.ptr(Dc),
                               O.ptr(Dp))
    r = H.cells()
    O.lib().o_hier_conc_rhs(H.h, C.byref(p), O.ptr(c), O.ptr(ph), O.ptr(Dc), O.ptr(Dp), O.ptr(r))
    w = H.weights()
    scale = np.abs(w * r).sum()
    assert abs((w * r).sum()) < 1e-13 * scale
    # the same levels evaluated one by one (no coarsened fine fluxes)
    r2 = H.cells()
    for pp in range(H.np):
        l = H.patch_level[pp]
        pl = P.two_grain_params()
        pl.dx[:] = [1.0 / (1 << l)] * 3
        G = H.geoms[pp]
        o, fo = H.coff[pp], H.foff[pp]
        sl = slice(o, o + G.ncg)
        blk = np.zeros(G.ncg)
        O.lib().o_conc_rhs(C.byref(G), C.byref(pl), O.ptr(np.ascontiguousarray(c[sl])),
                           O.ptr(np.ascontiguousarray(ph[sl])), O.ptr(Dc[fo:]), O.ptr(Dp[fo:]),
                           O.ptr(blk))
        r2[sl] = blk
    assert abs((w * r2).sum()) > 1e-6 * scale


def test_full_fine_level_reduces_to_single_level():
    """Level 1 covering the periodic domain: its composite RHS is the fine
    single-level RHS of the single-patch oracle."""
    lev1 = [((0, 0), (31, 63)), ((32, 0), (63, 63))]
    H = O.Hier(2, [LEV0, lev1], (32, 32), periodic=True)
    p = P.two_grain_params()
    p.flux_type = P.FLUX_ANISOTROPIC
    p.aniso_nu = 0.05
    p.aniso_knumber = 4
    rng = np.random.default_rng(5)
    phi1 = rng.uniform(0, 1, (64, 64))
    q1 = rng.normal(size=(4, 64, 64))
    q1 /= np.sqrt((q1 ** 2).sum(0))
    phi = H.cells()
    q = H.cells(4)
    H.set_level(phi, 1, phi1)
    for m in range(4):
        H.set_level(q, 1, q1[m], comp=m)
    O.lib().o_hier_coarsen(H.h, O.ptr(phi), 1, 0)
    O.lib().o_hier_coarsen(H.h, O.ptr(q), 4, 1)
    O.lib().o_hier_ghost_fill(H.h, O.ptr(phi), 1, 2)
    O.lib().o_hier_ghost_fill(H.h, O.ptr(q), 4, 2)
    r = H.cells()
    p.with_conc = 0
    O.lib().o_hier_phase_rhs(H.h, C.byref(p), O.ptr(phi), O.ptr(q), None, None, O.ptr(r))
    got = H.get_level(r, 1)
    # single patch, fine spacing 1/2
    pf = P.two_grain_params()
    pf.flux_type, pf.aniso_nu, pf.aniso_knumber, pf.with_conc = P.FLUX_ANISOTROPIC, 0.05, 4, 0
    pf.dx[:] = [0.5, 0.5, 0.5]
    pt = O.Patch(2, (64, 64, 1), g=1, dx=(0.5, 0.5, 1.0))
    a = pt.cells()
    pt.interior(a)[...] = phi1
    qa = pt.cells(4)
    pt.interior(qa)[...] = q1.reshape(pt.interior(qa).shape)
    O.lib().o_ghost_fill(C.byref(pt.G), O.ptr(a), 1)
    O.lib().o_ghost_fill(C.byref(pt.G), O.ptr(qa), 4)
    ro = pt.cells()
    O.lib().o_phase_rhs(C.byref(pt.G), C.byref(pf), O.ptr(a), O.ptr(qa), None, None, O.ptr(ro))
    ref = pt.interior(ro).reshape(64, 64)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("ewing", [1, 0])
def test_ewing_fix_exact_on_linear_fields(ewing):
    H = O.Hier(2, [LEV0, LEV1], (32, 32), periodic=False)
    u = _fill_levels(H, lambda l, n: _lin(n, 1.0 / (1 << l), 0.0, 1.0, 0.5))
    f = H.cells()
    B = O.HierBlock(kind=0, D=None, D_const=-1.0, C=None, C_const=0.0, M=None, M_const=1.0,
                    S=None, gamma=0.0, ewing=ewing)
    r = H.cells()
    O.lib().o_hier_comp_residual(H.h, C.byref(B), O.ptr(u), O.ptr(f), O.ptr(r), 1)
    # fine cells next to the coarse-fine boundary, away from the domain boundary
    res = H.get_level(r, 1, fill=0.0)
    inner = np.abs(res[2:-2, 2:-2])
    coarse = np.abs(H.get_level(r, 0, fill=0.0)[2:-2, 2:-2] * (H.uncovered(0)[2:-2, 2:-2]))
    if ewing:
        assert inner.max() < 1e-6 and coarse.max() < 1e-6, (inner.max(), coarse.max())
    else:
        assert inner.max() > 1e-2


def _fac(H, B, f, cycles, ncomp=1):
    FP = O.HierFacParams(max_cycles=cycles, pre=1, post=1, coarse_its=2, coarse_tol=-1.0)
    u = H.cells(ncomp)
    nrm, cyc = C.c_double(), C.c_int()
    O.lib().o_hier_fac_solve(H.h, C.byref(B), C.byref(FP), O.ptr(f), O.ptr(u), ncomp, 1e-30,
                             C.byref(nrm), C.byref(cyc))
    return u, nrm.value


def test_fac_converges_and_matches_fine_solution():
    gamma = 1e-2
    B = O.HierBlock(kind=0, D=None, D_const=-gamma * 2.0, C=None, C_const=1.0, M=None,
                    M_const=1.0, S=None, gamma=0.0, ewing=1)
    H = O.Hier(2, [LEV0, LEV1], (32, 32), periodic=True)
    rng = np.random.default_rng(7)
    f = _fill_levels(H, lambda l, n: rng.standard_normal((n, n)))
    norms = [_fac(H, B, f, k)[1] for k in (1, 2, 3, 4)]
    for a, b in zip(norms[:-1], norms[1:]):
        assert b < 0.1 * a
    # fine level over the whole domain: the fine single-level solution
    lev1 = [((0, 0), (31, 63)), ((32, 0), (63, 63))]
    H2 = O.Hier(2, [LEV0, lev1], (32, 32), periodic=True)
    f1 = rng.standard_normal((64, 64))
    f2 = H2.cells()
    H2.set_level(f2, 1, f1)
    u, nrm = _fac(H2, B, f2, 12)
    assert nrm < 1e-11
    pt = O.Patch(2, (64, 64, 1), g=1, dx=(0.5, 0.5, 1.0))
    op = O.EllipticOp()
    op.D, op.D_const, op.C, op.C_const, op.M, op.M_const = None, -gamma * 2.0, None, 1.0, None, 1.0
    op.dx = (C.c_double * 3)(0.5, 0.5, 1.0)
    rhs = pt.cells()
    pt.interior(rhs)[...] = f1
    us = pt.cells()
    flux = np.zeros(pt.nf)
    O.lib().o_fac_smooth(C.byref(pt.G), C.byref(op), O.ptr(rhs), O.ptr(us), 400, -1.0, O.ptr(flux))
    ref = pt.interior(us).reshape(64, 64)
    assert np.max(np.abs(H2.get_level(u, 1) - ref)) < 1e-9 * np.max(np.abs(ref))


def _two_grain():
    p = P.two_grain_params()
    p.quad_ceq[0], p.quad_ceq[1] = 0.40, 0.45
    return p


def test_hier_integrator_single_level_is_the_patch_integrator():
    """One level covering the periodic domain: the hierarchy integrator takes
    the single-patch integrator's steps (same counters, fields to rounding)."""
    n = 32
    p = _two_grain()
    s = P.solver_defaults(maxl=10, max_newton=5)
    phi, c, q = P.two_grain_state(n=n, radius=n / 4)
    H = O.Hier(2, [[((0, 0), (n - 1, n - 1))]], (n, n), tuple(p.dx), True)
    y = H.cells(2 + len(q))
    for k, f in enumerate([phi, c] + list(q)):
        H.set_level(y, 0, np.asarray(f, dtype=np.float64), comp=k)
    hs = O.HierSim(p, s, H)
    hs.set_state(y)
    ss = O.Sim(p, s, n, True)
    ss.set_state(phi, c, np.asarray(q))
    for _ in range(3):
        r1, a = hs.advance(1)
        r2, b = ss.advance(1)
        assert r1 == r2 == 0
        for k in ("nst", "nni", "nli", "nps", "ncfn", "nlcf"):
            assert a[k] == b[k], (k, a, b)
        assert a["t"] == b["t"]
        yh = hs.get_state()
        ph, cc, qq, _, _ = ss.get_state()
        assert np.abs(H.get_level(yh, 0, comp=0) - ph).max() < 1e-14
        assert np.abs(H.get_level(yh, 0, comp=1) - cc).max() < 1e-14
        for m in range(len(q)):
            assert np.abs(H.get_level(yh, 0, comp=2 + m) - qq[m]).max() < 1e-14


def test_hier_integrator_full_fine_level_is_the_fine_integrator():
    """Level 1 covering the periodic domain: the two-level integrator follows
    the single-patch integrator on the fine grid (the coarse level only
    changes the preconditioner, so tight tolerances give the same steps and
    fields within 1e-10)."""
    n = 16
    p = _two_grain()
    pf = _two_grain()
    pf.dx[:] = [d / 2 for d in p.dx]
    s = P.solver_defaults(maxl=20, max_newton=8, atol=1e-10, rtol=1e-12)
    phi, c, q = P.two_grain_state(n=2 * n, radius=n / 2)
    fields = [np.asarray(f, dtype=np.float64) for f in [phi, c] + list(q)]
    H = O.Hier(2, [[((0, 0), (n - 1, n - 1))], [((0, 0), (2 * n - 1, 2 * n - 1))]], (n, n),
               tuple(p.dx), True)
    y = H.cells(len(fields))
    for k, f in enumerate(fields):
        H.set_level(y, 1, f, comp=k)
        H.set_level(y, 0, 0.25 * (f[0::2, 0::2] + f[1::2, 0::2] + f[0::2, 1::2] + f[1::2, 1::2]),
                    comp=k)
    hs = O.HierSim(p, s, H)
    hs.set_state(y)
    ss = O.Sim(pf, s, 2 * n, True)
    ss.set_state(phi, c, np.asarray(q))
    for _ in range(3):
        r1, a = hs.advance(1)
        r2, b = ss.advance(1)
        assert r1 == r2 == 0
        assert a["nst"] == b["nst"] and a["nni"] == b["nni"] and a["t"] == b["t"], (a, b)
        yh = hs.get_state()
        ph, cc, qq, _, _ = ss.get_state()
        assert np.abs(H.get_level(yh, 1, comp=0) - ph).max() < 1e-10
        assert np.abs(H.get_level(yh, 1, comp=1) - cc).max() < 1e-10
        for m in range(len(q)):
            assert np.abs(H.get_level(yh, 1, comp=2 + m) - qq[m]).max() < 1e-10


def test_transfer_copies_old_cells_and_refines_new_ones():
    """o_hier_transfer: identity onto the same hierarchy; a new fine patch
    takes the old fine values where the old level holds the cell and the
    exact linear values elsewhere on a linear field (linear refinement is
    exact away from the clamped boundary); a non-linear old fine value
    survives the transfer."""
    Ho = O.Hier(2, [LEV0, LEV1], (32, 32), periodic=False)
    f = _fill_levels(Ho, lambda l, n: _lin(n, 1.0 / (1 << l)))
    # mark one old fine cell with a value linear refinement cannot produce
    lv1 = Ho.get_level(f, 1, fill=0.0)
    lv1[20, 18] = 7.0
    Ho.set_level(f, 1, lv1)
    same = O.hier_transfer(Ho, Ho, f)
    for l in range(2):
        a, b = Ho.get_level(same, l), Ho.get_level(f, l)
        m = ~np.isnan(b)
        assert np.array_equal(a[m], b[m])
    new1 = [((12, 14), (27, 29)), ((40, 40), (51, 55))]
    Hn = O.Hier(2, [LEV0, new1], (32, 32), periodic=False)
    g = O.hier_transfer(Ho, Hn, f)
    got = Hn.get_level(g, 1)
    exact = _lin(64, 0.5)
    old = ~np.isnan(Ho.get_level(f, 1))
    for (lo, hi) in new1:
        for j in range(lo[1], hi[1] + 1):
            for i in range(lo[0], hi[0] + 1):
                want = lv1[j, i] if old[j, i] else exact[j, i]
                assert abs(got[j, i] - want) < 1e-14, (i, j)
    assert got[20, 18] == 7.0
    assert np.array_equal(Hn.get_level(g, 0), Ho.get_level(f, 0))


def _disk_phi(n, r, w):
    y, x = np.mgrid[0:n, 0:n] + 0.5
    d = np.sqrt((x - n / 2) ** 2 + (y - n / 2) ** 2)
    return 0.5 * (1.0 - np.tanh((d - r) / (w / 2.0)))


@pytest.mark.parametrize("periodic", [True, False])
def test_regrid_covers_tags_and_nests(periodic):
    """regrid_levels: every buffered tag of level k lies in a level-(k+1)
    box, boxes are disjoint and properly nested (the oracle hierarchy accepts
    them), level-k+1 boxes of k >= 1 stay one cell inside their level-k
    patch, and a field without gradients keeps level 0 alone."""
    n = 32
    Ho = O.Hier(2, [[((0, 0), (19, 31)), ((20, 0), (31, 31))]], (n, n), periodic=periodic)
    f = Ho.cells()
    Ho.set_level(f, 0, _disk_phi(n, 6.0, 3.0))
    p = P.two_grain_params()
    tol, buf = 0.05, 1
    levels = O.regrid_levels(Ho, f, p, tol, 0.7, (2, 2, 2), (8, 8, 8), buf, 3)
    assert len(levels) == 3
    Hn = O.Hier(2, levels, (n, n), periodic=periodic)  # nesting checked by the constructor
    O.lib().o_hier_ghost_fill(Ho.h, O.ptr(f), 1, 2)
    tags0 = O.hier_level_tags(Ho, f, 0, p, tol)
    cov = np.zeros((2 * n, 2 * n), bool)
    for (lo, hi) in levels[1]:
        assert not cov[lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].any()
        cov[lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    for j, i in zip(*np.nonzero(tags0 > 0)):
        for dj in range(-buf, buf + 1):
            for di in range(-buf, buf + 1):
                jj, ii = j + dj, i + di
                if 0 <= jj < n and 0 <= ii < n:
                    assert cov[2 * jj, 2 * ii], (ii, jj)
    for (lo, hi) in levels[2]:
        clo, chi = [x // 2 for x in lo], [x // 2 for x in hi]
        inside = [b for b in levels[1]
                  if all(b[0][a] + 1 <= clo[a] and chi[a] <= b[1][a] - 1 for a in range(2))]
        assert inside, (lo, hi)
    # the fine levels follow the field: the transferred disk tags level 1
    g = O.hier_transfer(Ho, Hn, f)
    O.lib().o_hier_ghost_fill(Hn.h, O.ptr(g), 1, 2)
    assert (O.hier_level_tags(Hn, g, 1, p, tol) > 0).any()
    flat = Ho.cells() + 0.25
    assert len(O.regrid_levels(Ho, flat, p, tol, 0.7, (2, 2, 2), (8, 8, 8), buf, 3)) == 1
