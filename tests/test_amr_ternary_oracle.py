"""Pins of the oracle's three-phase ternary multi-order model on a patch
hierarchy (o_hier_phase_rhs_tern3, o_hier_ebs3_coeffs, o_hier_conc_rhs_ebs3
and the hierarchy integrator's ternary branches) that do not retype them:

- composite conservation of both species: with every order's and species'
  fine fluxes coarsened onto the coarse-fine faces the composition
  right-hand sides integrate to zero on a periodic hierarchy;
- a fine level covering the whole domain reduces the composite phase and
  composition right-hand sides to the single-patch ones on the fine grid;
- on the Sintering deck's initial state, a single level covering the domain
  makes the hierarchy integrator the single-patch one (same counters), and a
  fine level covering the domain over a coarse one follows the fine
  single-patch integrator step by step within the Newton tolerance (the
  coarse level only enters the preconditioner).
"""
import ctypes as C

import numpy as np

from ampe_amd import params as P
from oracle import oracle as O

LEV0 = [((0, 0), (15, 31)), ((16, 0), (31, 31))]
LEV1 = [((16, 16), (31, 39)), ((32, 20), (47, 35))]


def _deck():
    p, (phi, c), chk = P.sintering_deck()
    return p, phi, c


def _random_fields(H, rng, norder):
    phi = H.cells(norder)
    ph6 = H.cells(6)
    for lev in range(len(H.levels)):
        n = H.dom0[0] << lev
        v = rng.uniform(0.05, 1.0, (norder, n, n))
        for m in range(norder):
            H.set_level(phi, lev, v[m], comp=m)
        for m in range(6):
            H.set_level(ph6, lev, rng.uniform(0.05, 0.5, (n, n)), comp=m)
    O.lib().o_hier_coarsen(H.h, O.ptr(phi), norder, 0)
    O.lib().o_hier_coarsen(H.h, O.ptr(ph6), 6, 0)
    O.lib().o_hier_ghost_fill(H.h, O.ptr(phi), norder, 2)
    O.lib().o_hier_ghost_fill(H.h, O.ptr(ph6), 6, 2)
    return phi, ph6


def test_composite_conservation_of_both_species():
    p, _, _ = _deck()
    H = O.Hier(2, [LEV0, LEV1], (32, 32), dx0=(0.05, 0.05, 1.0), periodic=True)
    rng = np.random.default_rng(8)
    phi, ph6 = _random_fields(H, rng, p.norderp)
    DL, DA, DB = H.sides(), H.sides(), H.sides()
    O.lib().o_hier_ebs3_coeffs(H.h, C.byref(p), p.norderp, O.ptr(phi), O.ptr(DL), O.ptr(DA),
                               O.ptr(DB))
    assert np.abs(DA).max() > 0 and np.abs(DB).max() > 0
    r = H.cells(2)
    O.lib().o_hier_conc_rhs_ebs3(H.h, C.byref(p), O.ptr(ph6), O.ptr(DL), O.ptr(DA), O.ptr(DB),
                                 O.ptr(r))
    w = H.weights()
    for k in range(2):
        rk = r[k * H.ncg:(k + 1) * H.ncg]
        assert abs((w * rk).sum()) < 1e-12 * np.abs(w * rk).sum()


def test_full_fine_level_reduces_to_single_level():
    p, _, _ = _deck()
    n = 16
    lev1 = [((0, 0), (2 * n - 1, 2 * n - 1))]
    H = O.Hier(2, [[((0, 0), (n - 1, n - 1))], lev1], (n, n), dx0=(0.1, 0.1, 1.0),
               periodic=True)
    rng = np.random.default_rng(4)
    no = p.norderp
    v = rng.uniform(0.05, 1.0, (no, 2 * n, 2 * n))
    cv = rng.uniform(0.05, 0.5, (6, 2 * n, 2 * n))
    phi, ph6 = H.cells(no), H.cells(6)
    for m in range(no):
        H.set_level(phi, 1, v[m], comp=m)
    for m in range(6):
        H.set_level(ph6, 1, cv[m], comp=m)
    O.lib().o_hier_coarsen(H.h, O.ptr(phi), no, 0)
    O.lib().o_hier_coarsen(H.h, O.ptr(ph6), 6, 0)
    O.lib().o_hier_ghost_fill(H.h, O.ptr(phi), no, 2)
    O.lib().o_hier_ghost_fill(H.h, O.ptr(ph6), 6, 2)
    r = H.cells(no)
    O.lib().o_hier_phase_rhs_tern3(H.h, C.byref(p), no, O.ptr(phi), O.ptr(ph6), O.ptr(r))
    DL, DA, DB = H.sides(), H.sides(), H.sides()
    O.lib().o_hier_ebs3_coeffs(H.h, C.byref(p), no, O.ptr(phi), O.ptr(DL), O.ptr(DA), O.ptr(DB))
    rc = H.cells(2)
    O.lib().o_hier_conc_rhs_ebs3(H.h, C.byref(p), O.ptr(ph6), O.ptr(DL), O.ptr(DA), O.ptr(DB),
                                 O.ptr(rc))
    # the single patch on the fine grid
    pf = P.sintering_deck()[0]
    pf.dx[:] = [0.05, 0.05, 1.0]
    pt = O.Patch(2, (2 * n, 2 * n, 1), g=1, dx=(0.05, 0.05, 1.0))
    a, b = pt.cells(no), pt.cells(6)
    pt.interior(a)[...] = v[:, None]
    pt.interior(b)[...] = cv[:, None]
    O.lib().o_ghost_fill(C.byref(pt.G), O.ptr(a), no)
    O.lib().o_ghost_fill(C.byref(pt.G), O.ptr(b), 6)
    ro = pt.cells(no)
    O.lib().o_phase_rhs_tern3(C.byref(pt.G), C.byref(pf), no, O.ptr(a), O.ptr(b), None,
                              O.ptr(ro))
    dl, da, db = (np.zeros(pt.nf) for _ in range(3))
    O.lib().o_ebs3_coeffs(C.byref(pt.G), C.byref(pf), no, O.ptr(a), O.ptr(dl), O.ptr(da),
                          O.ptr(db))
    rco = pt.cells(2)
    O.lib().o_conc_rhs_ebs3(C.byref(pt.G), C.byref(pf), O.ptr(b), O.ptr(dl), O.ptr(da),
                            O.ptr(db), O.ptr(rco))
    for m in range(no):
        got = H.get_level(r, 1, comp=m)
        ref = pt.interior(ro)[m, 0]
        assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref)), m
    for k in range(2):
        got = H.get_level(rc, 1, comp=k)
        ref = pt.interior(rco)[k, 0]
        assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref)), k


def _hier_state(H, fields, lev):
    y = H.cells(len(fields))
    for k, f in enumerate(fields):
        H.set_level(y, lev, f, comp=k)
        if lev == 1:
            H.set_level(y, 0, 0.25 * (f[0::2, 0::2] + f[1::2, 0::2] + f[0::2, 1::2] +
                                      f[1::2, 1::2]), comp=k)
    return y


def test_hier_integrator_single_level_is_the_patch_integrator():
    """One level covering the Sintering deck's periodic 64^2 domain: the
    hierarchy integrator takes the single-patch integrator's steps, Newton
    failures and step-size cuts included (tight tolerances force one)."""
    p, phi, c = _deck()
    n = 64
    s = P.solver_defaults(dt=1e-5, maxl=20, max_newton=8, atol=1e-8, rtol=1e-10)
    H = O.Hier(2, [[((0, 0), (n - 1, n - 1))]], (n, n), tuple(p.dx), True)
    hs = O.HierSim(p, s, H)
    hs.set_state(_hier_state(H, list(phi) + list(c), 0))
    ss = O.Sim(p, s, n, True)
    ss.set_state(phi, c, None)
    for _ in range(3):
        r1, a = hs.advance(1)
        r2, b = ss.advance(1)
        assert r1 == r2 == 0
        for k in ("nst", "nni", "nli", "nps", "ncfn", "nlcf"):
            assert a[k] == b[k], (k, a, b)
        assert a["t"] == b["t"]
        yh = hs.get_state()
        ph, cc = ss.get_state()[:2]
        for m in range(p.norderp):
            assert np.abs(H.get_level(yh, 0, comp=m) - ph[m]).max() < 1e-13
        for k in range(2):
            assert np.abs(H.get_level(yh, 0, comp=p.norderp + k) - cc[k]).max() < 1e-12
    assert b["ncfn"] >= 1


def test_hier_integrator_full_fine_level_is_the_fine_integrator():
    """The Sintering deck's 64^2 initial state as a fine level covering the
    periodic domain over a 32^2 coarse level: the coarse level only changes
    the preconditioner, so the two integrators take the same steps and
    Newton iterations and their fields differ by the Newton tolerance
    (atol 1e-7 here: within 1e-6)."""
    p, phi, c = _deck()
    n = 32
    pc = P.sintering_deck()[0]
    pc.dx[:] = [2 * d for d in p.dx]
    s = P.solver_defaults(dt=1e-5, maxl=40, max_newton=8, atol=1e-7, rtol=1e-9)
    H = O.Hier(2, [[((0, 0), (n - 1, n - 1))], [((0, 0), (2 * n - 1, 2 * n - 1))]], (n, n),
               tuple(pc.dx), True)
    hs = O.HierSim(pc, s, H)
    hs.set_state(_hier_state(H, list(phi) + list(c), 1))
    ss = O.Sim(p, s, 2 * n, True)
    ss.set_state(phi, c, None)
    for _ in range(3):
        r1, a = hs.advance(1)
        r2, b = ss.advance(1)
        assert r1 == r2 == 0
        assert a["nst"] == b["nst"] and a["nni"] == b["nni"] and a["t"] == b["t"], (a, b)
        # the two-level preconditioner is the stronger one
        assert a["nli"] <= b["nli"]
        yh = hs.get_state()
        ph, cc = ss.get_state()[:2]
        for m in range(p.norderp):
            assert np.abs(H.get_level(yh, 1, comp=m) - ph[m]).max() < 1e-6
        for k in range(2):
            assert np.abs(H.get_level(yh, 1, comp=p.norderp + k) - cc[k]).max() < 1e-6


# ---- the multi-order model without composition (grain growth) on a hierarchy

def _grains_params():
    p = P.two_grain_params(ndim=2)
    p.norderp = 4
    p.with_conc = 0
    p.evolve_quat = 0
    p.flux_type = P.FLUX_SIMPLE
    return p


def _grains_state(n, seed=3):
    """Four smoothed Voronoi grains on a periodic n^2 grid."""
    rng = np.random.default_rng(seed)
    seeds = rng.uniform(0, n, (4, 2))
    x = np.arange(n) + 0.5
    yy, xx = np.meshgrid(x, x, indexing="ij")
    d = []
    for sx, sy in seeds:
        dx = np.abs(xx - sx)
        dy = np.abs(yy - sy)
        dx, dy = np.minimum(dx, n - dx), np.minimum(dy, n - dy)
        d.append(np.sqrt(dx * dx + dy * dy))
    d = np.array(d)
    w = np.exp(-(d - d.min(0)) / 1.5)
    return list(w / w.sum(0))


def test_multiorder_flux_part_is_conservative():
    """With no wells (phi_well_scale = 0) the composite multi-order RHS is
    M div(eps^2/2 grad phi_p): it integrates to zero over a periodic
    hierarchy for every order, which needs the fine fluxes on the
    coarse-fine faces."""
    p = _grains_params()
    p.phi_well_scale = 0.0
    H = O.Hier(2, [LEV0, LEV1], (32, 32), dx0=(0.05, 0.05, 1.0), periodic=True)
    phi, _ = _random_fields(H, np.random.default_rng(5), p.norderp)
    r = H.cells(p.norderp)
    O.lib().o_hier_phase_rhs_multiorder(H.h, C.byref(p), p.norderp, O.ptr(phi), O.ptr(r))
    w = H.weights()
    for m in range(p.norderp):
        rm = r[m * H.ncg:(m + 1) * H.ncg]
        assert np.abs(w * rm).sum() > 0
        assert abs((w * rm).sum()) < 1e-12 * np.abs(w * rm).sum()


def test_multiorder_hier_single_level_is_the_patch_integrator():
    p = _grains_params()
    n = 32
    s = P.solver_defaults(dt=1e-3, maxl=10, max_newton=5, atol=1e-8, rtol=1e-10)
    phi = _grains_state(n)
    H = O.Hier(2, [[((0, 0), (n - 1, n - 1))]], (n, n), tuple(p.dx), True)
    hs = O.HierSim(p, s, H)
    hs.set_state(_hier_state(H, phi, 0))
    ss = O.Sim(p, s, n, True)
    ss.set_state(phi, None, None)
    for _ in range(3):
        r1, a = hs.advance(1)
        r2, b = ss.advance(1)
        assert r1 == r2 == 0
        for k in ("nst", "nni", "nli", "nps", "ncfn", "nlcf"):
            assert a[k] == b[k], (k, a, b)
        yh = hs.get_state()
        ph = ss.get_state()[0]
        for m in range(p.norderp):
            assert np.abs(H.get_level(yh, 0, comp=m) - ph[m]).max() < 1e-13
