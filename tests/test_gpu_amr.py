"""GPU parity of the hierarchy (multi-level) operators against the oracle's
restatement (oracle/oracle_amr.c): coarse-fine ghost fill by constant and
linear refinement, conservative coarsening, composite right-hand sides with
the fine fluxes coarsened onto the coarse-fine faces, and the FAC
preconditioner cycled over levels with Ewing's flux fix.

Hierarchies: 2-D periodic and bounded two-level (uneven fine patches, one
touching the physical boundary, two meeting inside the fine level), a 2-D
three-level one, and 3-D periodic and bounded two-level ones.  Composite
fields compare on the cells not covered by a finer level (their values are
the ones the reference keeps; covered cells are overwritten by coarsening).
Tolerances: 1e-15 for transfers (same operations), 1e-13 / 1e-12 (with
transcendentals) for right-hand sides, 1e-12 for FAC solutions.
"""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu

HIERS = {
    "2d_periodic": (2, [[((0, 0), (15, 31)), ((16, 0), (31, 31))],
                        [((16, 16), (31, 39)), ((32, 20), (47, 35))]], (32, 32), True),
    "2d_bounded": (2, [[((0, 0), (15, 31)), ((16, 0), (31, 31))],
                       [((0, 0), (15, 11)), ((16, 16), (31, 39)), ((32, 20), (47, 35))]],
                   (32, 32), False),
    "2d_three_levels": (2, [[((0, 0), (15, 15))], [((8, 8), (23, 23)), ((4, 24), (11, 29))],
                            [((20, 20), (35, 35))]], (16, 16), False),
    "3d_periodic": (3, [[((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 15, 15))],
                        [((8, 8, 8), (19, 15, 15)), ((20, 8, 8), (27, 19, 13))]],
                    (16, 16, 16), True),
    "3d_bounded": (3, [[((0, 0, 0), (15, 15, 15))],
                       [((0, 0, 0), (7, 7, 7)), ((12, 12, 12), (23, 21, 19))]],
                   (16, 16, 16), False),
}


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Pair:
    """The oracle hierarchy and the device hierarchy over the same boxes;
    composite cell arrays share one layout, side fields differ and are
    mapped here."""

    def __init__(self, name, dev, dx0=1.0):
        from oracle import oracle as O
        from ampe_amd import hip as H
        self.O, self.H, self.dev = O, H, dev
        ndim, levels, dom0, periodic = HIERS[name]
        self.ndim = ndim
        dx = (dx0, dx0, dx0)
        self.ho = O.Hier(ndim, levels, dom0, dx, periodic)
        self.hg = H.Hier(ndim, levels, dom0, dx, periodic)
        assert self.ho.ncg == self.hg.ncg
        self.ncg = self.ho.ncg
        self.levels = levels
        self.dom = list(dom0) + ([1] if ndim == 2 else [])
        # side-field maps: oracle index -> device index per face
        o2g = np.empty(self.ho.nf, dtype=np.int64)
        soff = [0, 0, 0]
        per_dir = []
        for p, (lo, hi) in enumerate(self.ho.boxes):
            n = [hi[a] - lo[a] + 1 for a in range(ndim)] + ([1] if ndim == 2 else [])
            G = self.ho.geoms[p]
            o = self.ho.foff[p]
            blocks = []
            for d in range(ndim):
                cnt = (n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2))
                blocks.append((d, o, soff[d], cnt))
                o += cnt
                soff[d] += cnt
            per_dir.append(blocks)
        dir_off = [0, soff[0], soff[0] + soff[1]]
        for blocks in per_dir:
            for d, o, s, cnt in blocks:
                o2g[o:o + cnt] = dir_off[d] + s + np.arange(cnt)
        assert soff[:ndim] == self.hg.nside[:ndim]
        self.o2g = o2g
        self.w = self.ho.weights()
        self.unc = self.w > 0

    def dev_cells(self, a):
        import torch
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.dev)

    def dev_sides(self, a, ncomp=1):
        import torch
        out = np.zeros(ncomp * self.hg.nside_tot)
        for m in range(ncomp):
            out[m * self.hg.nside_tot + self.o2g] = a[m * self.ho.nf:(m + 1) * self.ho.nf]
        return torch.from_numpy(out).to(self.dev)

    def host_sides(self, t, ncomp=1):
        h = t.cpu().numpy()
        out = np.zeros(ncomp * self.ho.nf)
        for m in range(ncomp):
            out[m * self.ho.nf:(m + 1) * self.ho.nf] = h[m * self.hg.nside_tot + self.o2g]
        return out

    def field(self, seed, lo=0.0, hi=1.0, ncomp=1, unit=False):
        """Random values on every level's cells (coarsened, then the
        ghosts filled by linear refinement on both sides)."""
        rng = _rng(seed)
        a = self.ho.cells(ncomp)
        for lev in range(len(self.levels)):
            n = [s << lev for s in self.dom[:self.ndim]]
            shape = tuple(reversed(n))
            vals = rng.uniform(lo, hi, (ncomp,) + shape)
            if unit:
                vals /= np.sqrt((vals ** 2).sum(0))
            for m in range(ncomp):
                self.ho.set_level(a, lev, vals[m], comp=m)
        O = self.O
        O.lib().o_hier_coarsen(self.ho.h, O.ptr(a), ncomp, 1 if unit else 0)
        O.lib().o_hier_ghost_fill(self.ho.h, O.ptr(a), ncomp, 2)
        return a

    def interior_mask(self, ncomp=1):
        m = np.zeros(self.ncg, dtype=bool)
        for p, (lo, hi) in enumerate(self.ho.boxes):
            v = self.ho.ghost_view(m.view(np.uint8), p)
            if self.ndim == 3:
                v[1:-1, 1:-1, 1:-1] = 1
            else:
                v[1:-1, 1:-1] = 1
        return np.tile(m, ncomp)


CASES = list(HIERS)


@pytest.mark.parametrize("name", CASES)
def test_transfers(dev, name):
    """Weights, ghost fills (constant and linear refine) and conservative
    coarsening (plain and normalised) give the oracle's values."""
    import torch
    print(f"hier {name}: transfers", flush=True)
    X = Pair(name, dev)
    O, H = X.O, X.H
    wg = torch.zeros(X.ncg, dtype=torch.float64, device=dev)
    H.check(H.lib().ampe_hier_weights(X.hg.h, _p(wg), None))
    assert np.array_equal(wg.cpu().numpy(), X.w)
    a = X.field(3, ncomp=3)
    for cf in (1, 2):
        b = a.copy()
        O.lib().o_hier_ghost_fill(X.ho.h, O.ptr(b), 3, cf)
        t = X.dev_cells(a)
        H.check(H.lib().ampe_hier_ghost_fill(X.hg.h, _p(t), 3, cf, None))
        assert np.max(np.abs(t.cpu().numpy() - b)) < 1e-15, cf
    q = X.field(5, lo=-1.0, hi=1.0, ncomp=4)
    for qn in (0, 1):
        b = q.copy()
        O.lib().o_hier_coarsen(X.ho.h, O.ptr(b), 4, qn)
        t = X.dev_cells(q)
        H.check(H.lib().ampe_hier_coarsen(X.hg.h, _p(t), 4, qn, None))
        im = X.interior_mask(4)
        assert np.max(np.abs(t.cpu().numpy()[im] - b[im])) < 1e-15, qn


def _phase_params(ndim, kind, dx0):
    if kind == "calphad":
        p = P.polycrystal_params(ndim=ndim)
    else:
        p = P.two_grain_params(ndim=ndim)
        if kind == "iso":
            p.flux_type = P.FLUX_ISOTROPIC
        elif kind == "aniso":
            p.flux_type = P.FLUX_ANISOTROPIC
            p.aniso_nu = 0.05
            p.aniso_knumber = 4
    p.dx[:] = [dx0, dx0, dx0]
    return p


@pytest.mark.parametrize("name", CASES)
def test_composite_rhs(dev, name):
    """Composite phase, orientation and composition right-hand sides: the
    uncovered cells next to the coarse-fine boundary carry the coarsened
    fine fluxes."""
    import torch
    ndim = HIERS[name][0]
    kinds = ["simple", "aniso", "calphad"] + (["iso"] if ndim == 2 else [])
    for kind in kinds:
        print(f"hier {name}: rhs {kind}", flush=True)
        dx0 = 0.02 if kind == "calphad" else 1.0
        X = Pair(name, dev, dx0)
        O, H = X.O, X.H
        p = _phase_params(ndim, kind, dx0)
        phi = X.field(11)
        q = X.field(13, lo=-1.0, hi=1.0, ncomp=4, unit=True)
        c = X.field(17, 0.3, 0.5)
        cl = X.field(19, 0.2, 0.6)
        cs = X.field(23, 0.2, 0.6)
        r_o = X.ho.cells()
        O.lib().o_hier_phase_rhs(X.ho.h, C.byref(p), O.ptr(phi), O.ptr(q), O.ptr(cl), O.ptr(cs),
                                 O.ptr(r_o))
        t = [X.dev_cells(v) for v in (phi, q, cl, cs)]
        r = torch.zeros(X.ncg, dtype=torch.float64, device=dev)
        H.check(H.lib().ampe_hier_phase_rhs(X.hg.h, C.byref(p), *[_p(v) for v in t], _p(r), None))
        e = _rel(r.cpu().numpy()[X.unc], r_o[X.unc])
        assert e < (1e-13 if kind == "simple" else 1e-12), (kind, e)
        if kind != "simple":
            continue
        # orientation: face norms, then the composite RHS from them
        n2_o = X.ho.sides()
        O.lib().o_hier_quat_face_norm2(X.ho.h, C.byref(p), O.ptr(q), O.ptr(n2_o))
        n2 = torch.zeros(X.hg.nside_tot, dtype=torch.float64, device=dev)
        H.check(H.lib().ampe_hier_quat_face_norm2(X.hg.h, C.byref(p), _p(t[1]), _p(n2), None))
        assert _rel(X.host_sides(n2), n2_o) < 1e-13
        rq_o = X.ho.cells(4)
        O.lib().o_hier_quat_rhs(X.ho.h, C.byref(p), O.ptr(phi), O.ptr(q), O.ptr(n2_o), O.ptr(rq_o))
        fc = torch.zeros_like(n2)
        rq = torch.zeros(4 * X.ncg, dtype=torch.float64, device=dev)
        H.check(H.lib().ampe_hier_quat_rhs(X.hg.h, C.byref(p), _p(t[0]), _p(t[1]), _p(n2), _p(fc),
                                           _p(rq), None))
        un4 = np.tile(X.unc, 4)
        e = _rel(rq.cpu().numpy()[un4], rq_o[un4])
        assert e < 1e-12, ("quat", e)
        # composition (KKS form): coefficients coarsened onto coarse faces
        Dc_o, Dp_o = X.ho.sides(), X.ho.sides()
        O.lib().o_hier_conc_coeffs(X.ho.h, C.byref(p), O.ptr(phi), O.ptr(cl), O.ptr(cs),
                                   O.ptr(Dc_o), O.ptr(Dp_o))
        Dc, Dp = torch.zeros_like(n2), torch.zeros_like(n2)
        H.check(H.lib().ampe_hier_conc_coeffs(X.hg.h, C.byref(p), _p(t[0]), _p(t[2]), _p(t[3]),
                                              _p(Dc), _p(Dp), None))
        assert _rel(X.host_sides(Dc), Dc_o) < 1e-12 and _rel(X.host_sides(Dp), Dp_o) < 1e-12
        rc_o = X.ho.cells()
        O.lib().o_hier_conc_rhs(X.ho.h, C.byref(p), O.ptr(c), O.ptr(phi), O.ptr(Dc_o),
                                O.ptr(Dp_o), O.ptr(rc_o))
        tc = X.dev_cells(c)
        rc = torch.zeros(X.ncg, dtype=torch.float64, device=dev)
        H.check(H.lib().ampe_hier_conc_rhs(X.hg.h, C.byref(p), _p(tc), _p(t[0]), _p(Dc), _p(Dp),
                                           _p(rc), None))
        e = _rel(rc.cpu().numpy()[X.unc], rc_o[X.unc])
        assert e < 1e-12, ("conc", e)
        # composite conservation of the composition flux (periodic domains)
        if HIERS[name][3]:
            s = (X.w * rc_o).sum()
            assert abs(s) < 1e-12 * np.abs(X.w * rc_o).sum()


@pytest.mark.parametrize("name", CASES)
def test_fac_solve(dev, name):
    """FAC cycles over the levels for the three block kinds: phase-like
    (C field, constant D), composition-like (D on faces, Ewing's fix) and
    the orientation block (four components, sqrt(M) scaling)."""
    import torch
    X = Pair(name, dev)
    O, H = X.O, X.H
    rng = _rng(31)
    gamma = 5e-3
    fp_o = O.HierFacParams(max_cycles=3, pre=1, post=1, coarse_its=2, coarse_tol=-1.0)
    fp_g = H.FacParams(max_cycles=3, pre=1, post=1, coarse_its=2, coarse_tol=-1.0)
    Cf = X.field(41, 1.0, 2.0)
    Dside = -gamma * rng.uniform(1.0, 5.0, X.ho.nf)
    S = X.field(43, 0.5, 1.5)
    fc = -rng.uniform(0.5, 2.0, X.ho.nf)
    for variant in ("phase", "conc", "quat"):
        print(f"hier {name}: fac {variant}", flush=True)
        ncomp = 4 if variant == "quat" else 1
        f = X.field(47, -1.0, 1.0, ncomp=ncomp)
        bo, bg = O.HierBlock(), H.HierBlock()
        keep = []
        if variant == "phase":
            bo.kind = bg.kind = 0
            bo.D = bg.D = None
            bo.D_const = bg.D_const = -gamma * 2.0
            tC = X.dev_cells(Cf)
            keep.append(tC)
            bo.C, bg.C = Cf.ctypes.data, tC.data_ptr()
            bo.M = bg.M = None
            bo.M_const = bg.M_const = 1.5
            bo.ewing = bg.ewing = 1
        elif variant == "conc":
            bo.kind = bg.kind = 0
            tD = X.dev_sides(Dside)
            keep.append(tD)
            bo.D, bg.D = Dside.ctypes.data, tD.data_ptr()
            bo.C = bg.C = None
            bo.C_const = bg.C_const = 1.0
            bo.M = bg.M = None
            bo.M_const = bg.M_const = 0.8
            bo.ewing = bg.ewing = 1
        else:
            bo.kind = bg.kind = 1
            tD = X.dev_sides(fc)
            tS = X.dev_cells(S)
            keep += [tD, tS]
            bo.D, bg.D = fc.ctypes.data, tD.data_ptr()
            bo.S, bg.S = S.ctypes.data, tS.data_ptr()
            bo.gamma = bg.gamma = gamma
            bo.ewing = bg.ewing = 0
        u_o = X.ho.cells(ncomp)
        nrm_o, cyc_o = C.c_double(), C.c_int()
        rc_o = O.lib().o_hier_fac_solve(X.ho.h, C.byref(bo), C.byref(fp_o), O.ptr(f), O.ptr(u_o),
                                        ncomp, 1e-30, C.byref(nrm_o), C.byref(cyc_o))
        tf = X.dev_cells(f)
        u = torch.zeros(ncomp * X.ncg, dtype=torch.float64, device=dev)
        nrm, cyc = C.c_double(), C.c_int()
        rc = H.lib().ampe_hier_fac_solve(X.hg.h, C.byref(bg), C.byref(fp_g), _p(tf), _p(u), ncomp,
                                         1e-30, C.byref(nrm), C.byref(cyc), None)
        assert rc == rc_o == 1 and cyc.value == cyc_o.value == 3, (rc, rc_o, cyc.value)
        im = X.interior_mask(ncomp)
        e = _rel(u.cpu().numpy()[im], u_o[im])
        print(f"  u rel err {e:.2e} norm {nrm.value:.3e} / {nrm_o.value:.3e}", flush=True)
        assert e < 1e-12, (variant, e)
        # the last norm is a small difference of large residual terms
        assert abs(nrm.value - nrm_o.value) < 1e-6 * nrm_o.value


def _refine2(a, periodic):
    """Cell-centred (bi/tri)linear refinement by 2 of a level array."""
    out = a
    for ax in range(a.ndim):
        n = out.shape[ax]
        idx = np.arange(n)
        up = (idx + 1) % n if periodic else np.minimum(idx + 1, n - 1)
        dn = (idx - 1) % n if periodic else np.maximum(idx - 1, 0)
        lo = 0.75 * out + 0.25 * np.take(out, dn, axis=ax)
        hi = 0.75 * out + 0.25 * np.take(out, up, axis=ax)
        sh = list(out.shape)
        sh[ax] = 2 * n
        r = np.empty(sh)
        sl_lo = [slice(None)] * out.ndim
        sl_hi = [slice(None)] * out.ndim
        sl_lo[ax] = slice(0, None, 2)
        sl_hi[ax] = slice(1, None, 2)
        r[tuple(sl_lo)] = lo
        r[tuple(sl_hi)] = hi
        out = r
    return out


def _hier_state(ho, fields, periodic):
    """Composite state: level 0 from the fields, finer levels by repeated
    linear refinement."""
    y = ho.cells(len(fields))
    for k, f in enumerate(fields):
        cur = np.asarray(f, dtype=np.float64)
        for lev in range(len(ho.levels)):
            ho.set_level(y, lev, cur, comp=k)
            cur = _refine2(cur, periodic)
    return y


def _disk_state(n, ndim, r, w, c0, seed):
    rng = _rng(seed)
    x = np.arange(n) + 0.5
    if ndim == 2:
        yy, xx = np.meshgrid(x, x, indexing="ij")
        d = np.sqrt((xx - n / 2) ** 2 + (yy - n / 2) ** 2)
        shape = (n, n)
    else:
        zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
        d = np.sqrt((xx - n / 2) ** 2 + (yy - n / 2) ** 2 + (zz - n / 2) ** 2)
        shape = (n, n, n)
    phi = 0.5 * (1.0 - np.tanh((d - r) / (w / 2.0)))
    phi = np.clip(phi + 0.01 * rng.standard_normal(shape) * (phi > 0.05) * (phi < 0.95), 0, 1)
    c = c0 + 0.002 * rng.uniform(-1, 1, shape)
    q = np.zeros((4,) + shape)
    q[0] = 1.0
    return phi, c, q


def _spheres3_state(n, nx=None):
    """Orders A0, A1, B and the liquid (1 - the others, clipped): spheres of
    radius n/4 with tanh interfaces on an nx x n x n grid (nx = n by
    default); c1 follows solid A, c2 solid B."""
    nx = nx or n
    x = np.arange(n) + 0.5
    zz, yy, xx = np.meshgrid(x, x, np.arange(nx) + 0.5, indexing="ij")
    phis = []
    for cx, cy in ((0.33, 0.35), (0.69, 0.35), (0.5, 0.68)):
        d = np.sqrt((xx - cx * nx) ** 2 + (yy - cy * n) ** 2 + (zz - n / 2) ** 2)
        phis.append(0.5 * (1.0 - np.tanh((d - n / 4) / 1.5)))
    liq = np.clip(1.0 - phis[0] - phis[1] - phis[2], 0.0, 1.0)
    c1 = 0.9 * (phis[0] + phis[1]) + 0.05
    c2 = 0.9 * phis[2] + 0.05
    return phis + [liq, c1, c2]


INTEGRATOR_CASES = {
    # two grains, quadratic KKS, orientation field, periodic, one fine box
    "2d_two_grain": dict(ndim=2, n=32, periodic=True,
                         levels=lambda n: [[((0, 0), (n - 1, n - 1))], [((16, 16), (47, 47))]],
                         steps=4),
    # dendrite-like: anisotropic flux (k = 4, nu = 0.05), dilute KKS with
    # k_e = 0.14, bounded domain, uneven fine patches
    "2d_aniso_dilute": dict(ndim=2, n=32, periodic=False,
                            levels=lambda n: [[((0, 0), (15, 31)), ((16, 0), (31, 31))],
                                              [((20, 20), (43, 35)), ((24, 36), (39, 43))]],
                            steps=4),
    # Sintering: three-phase ternary, five order parameters, two species,
    # periodic, two fine patches sharing an edge across the grain contacts
    "2d_sintering": dict(ndim=2, n=32, periodic=True,
                         levels=lambda n: [[((0, 0), (n - 1, n - 1))],
                                           [((16, 8), (47, 31)), ((20, 32), (43, 55))]],
                         steps=4),
    # the same with classical Gram-Schmidt (one pass for all of a Krylov
    # vector's weighted dots)
    "2d_sintering_cgs": dict(ndim=2, n=32, periodic=True,
                             levels=lambda n: [[((0, 0), (n - 1, n - 1))],
                                               [((16, 8), (47, 31)), ((20, 32), (43, 55))]],
                             steps=4),
    # grain growth: four order parameters, no composition or orientation
    "2d_grains": dict(ndim=2, n=32, periodic=True,
                      levels=lambda n: [[((0, 0), (n - 1, n - 1))],
                                        [((8, 12), (39, 35)), ((40, 12), (55, 51))]],
                      steps=4),
    # three-phase ternary in 3-D: two grains of solid A, one of B, the
    # liquid, one fine box over the three contacts
    "3d_ternary": dict(ndim=3, n=16, periodic=True,
                       levels=lambda n: [[((0, 0, 0), (n - 1, n - 1, n - 1))],
                                         [((6, 6, 8), (25, 25, 23))]],
                       steps=2),
    # the same wider than the 64-cell tiles: level 0 of 80 x 16 x 16 in two
    # uneven patches, a fine patch of 144 x 16 x 16 (three tiles along x)
    "3d_ternary_wide": dict(ndim=3, n=16, nx=80, periodic=True,
                            levels=lambda n: [[((0, 0, 0), (47, n - 1, n - 1)),
                                               ((48, 0, 0), (79, n - 1, n - 1))],
                                              [((8, 8, 8), (151, 23, 23))]],
                            steps=2),
    # 3-D CALPHAD polycrystal, one fine box
    "3d_calphad": dict(ndim=3, n=16, periodic=True,
                       levels=lambda n: [[((0, 0, 0), (n - 1, n - 1, n - 1))],
                                         [((8, 8, 8), (23, 23, 19))]],
                       steps=2),
}


def _integrator_case(case):
    """(ndim, n, periodic, levels, dom, model params, solver params, level-0
    fields) of an integrator case."""
    cfg = INTEGRATOR_CASES[case]
    ndim, n, periodic = cfg["ndim"], cfg["n"], cfg["periodic"]
    levels = cfg["levels"](n)
    if case == "2d_two_grain":
        p = P.two_grain_params()
        p.quad_ceq[0], p.quad_ceq[1] = 0.40, 0.45
        s = P.solver_defaults(maxl=10, max_newton=5)
        phi, c, q = P.two_grain_state(n=n, radius=n / 4)
        fields = [phi, c] + list(q)
    elif case == "2d_aniso_dilute":
        p = P.dilute_params(ndim=2, dilute_keq=0.14)
        p.flux_type = P.FLUX_ANISOTROPIC
        p.aniso_nu = 0.05
        p.aniso_knumber = 4
        s = P.solver_defaults(dt=1e-7, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8)
        phi, c, q = _disk_state(n, 2, 6.0, 3.0, 0.05, 7)
        fields = [phi, c] + list(q)
    elif case.startswith("2d_sintering"):
        p, (phi, c), _ = P.sintering_deck()
        p.dx[:] = [2 * d for d in p.dx]
        s = P.solver_defaults(dt=1e-5, maxl=14 if case.endswith("cgs") else 20, max_newton=8,
                              atol=1e-6, rtol=1e-8,
                              gstype=P.GS_CLASSICAL if case.endswith("cgs") else P.GS_MODIFIED)
        fields = [0.25 * (f[0::2, 0::2] + f[1::2, 0::2] + f[0::2, 1::2] + f[1::2, 1::2])
                  for f in list(phi) + list(c)]
    elif case == "2d_grains":
        p = P.two_grain_params(ndim=2)
        p.norderp = 4
        p.with_conc = 0
        p.evolve_quat = 0
        p.flux_type = P.FLUX_SIMPLE
        s = P.solver_defaults(dt=1e-3, maxl=10, max_newton=5, atol=1e-8, rtol=1e-10)
        rng = _rng(3)
        x = np.arange(n) + 0.5
        yy, xx = np.meshgrid(x, x, indexing="ij")
        d = []
        for sx, sy in rng.uniform(0, n, (4, 2)):
            ddx, ddy = np.abs(xx - sx), np.abs(yy - sy)
            ddx, ddy = np.minimum(ddx, n - ddx), np.minimum(ddy, n - ddy)
            d.append(np.sqrt(ddx * ddx + ddy * ddy))
        d = np.array(d)
        w = np.exp(-(d - d.min(0)) / 1.5)
        fields = list(w / w.sum(0))
    elif case.startswith("3d_ternary"):
        p = P.sintering_deck()[0]
        p.ndim = 3
        p.norderp, p.norderp_A = 4, 2
        s = P.solver_defaults(dt=1e-5, maxl=20, max_newton=8, atol=1e-6, rtol=1e-8)
        fields = _spheres3_state(n, cfg.get("nx"))
    else:
        p = P.polycrystal_params(ndim=3)
        s = P.solver_defaults(dt=1e-7, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8)
        phi, c, q = P.polycrystal_state(n=n, seed=1002, ngrains=4)
        fields = [phi, c] + list(q)
    dom = (cfg.get("nx", n),) + (n,) * (ndim - 1)
    return ndim, n, periodic, levels, dom, p, s, fields


@pytest.mark.parametrize("case", list(INTEGRATOR_CASES))
def test_hier_integrator_matches_oracle(dev, case):
    """BDF / Newton / SPGMR over the hierarchy on the device against the
    oracle, step by step: identical counters, composite fields (finer levels
    coarsened onto covered cells) within 1e-10."""
    from oracle import oracle as O
    from ampe_amd import hip as H
    ndim, n, periodic, levels, dom, p, s, fields = _integrator_case(case)
    cfg = INTEGRATOR_CASES[case]
    ho = O.Hier(ndim, levels, dom, tuple(p.dx), periodic)
    hg = H.Hier(ndim, levels, dom, tuple(p.dx), periodic)
    y0 = _hier_state(ho, fields, periodic)
    so = O.HierSim(p, s, ho)
    sg = H.HierSim(p, s, hg)
    so.set_state(y0)
    sg.set_state(y0)
    nc = so.nc
    im = np.zeros(ho.ncg, dtype=bool)
    for pp in range(ho.np):
        v = ho.ghost_view(im.view(np.uint8), pp)
        v[(slice(1, -1),) * ndim] = 1
    for step in range(cfg["steps"]):
        print(f"hier integrator {case}: step {step}", flush=True)
        ro, sto = so.advance(1)
        rg, stg = sg.advance(1)
        assert ro == rg == 0, (step, sto, stg)
        for k in ("nst", "nni", "nli", "ncfn", "nlcf", "npe", "nps"):
            assert sto[k] == stg[k], (step, k, sto, stg)
        assert abs(sto["t"] - stg["t"]) <= 1e-12 * sto["t"]
        yo, yg = so.get_state(), sg.get_state()
        for m in range(nc):
            a = yg[m * ho.ncg:(m + 1) * ho.ncg][im]
            b = yo[m * ho.ncg:(m + 1) * ho.ncg][im]
            e = _rel(a, b)
            assert e < 1e-10, (step, m, e)


@pytest.mark.parametrize("case", ["2d_two_grain", "2d_sintering", "2d_sintering_cgs", "2d_grains",
                                  "3d_ternary"])
def test_hier_fp32_preconditioner_matches_fp64(dev, case):
    """The FP32 FAC preconditioner over the hierarchy (every block solve in
    FP32, the Newton-Krylov outer loop in FP64) against the FP64 one on the
    device: with a Newton tolerance of 1e-10 both integrators converge to the
    same solution of each step's nonlinear system, so the composite fields
    agree to 1e-8 relative whatever the preconditioner's precision."""
    from ampe_amd import hip as H
    ndim, n, periodic, levels, dom, p, s, fields = _integrator_case(case)
    s.atol, s.rtol = 1e-10, 1e-12
    s.maxl, s.max_newton = 14, 10
    # the classical Gram-Schmidt case takes the Krylov path that fuses the
    # FP32 preconditioner input on one level (not on a hierarchy)
    s32 = P.SolverParams.from_buffer_copy(s)
    s32.precision = P.FAC_FP32
    hg = H.Hier(ndim, levels, dom, tuple(p.dx), periodic)
    from oracle import oracle as O
    ho = O.Hier(ndim, levels, dom, tuple(p.dx), periodic)
    y0 = _hier_state(ho, fields, periodic)
    a, b = H.HierSim(p, s, hg), H.HierSim(p, s32, hg)
    a.set_state(y0)
    b.set_state(y0)
    im = np.zeros(ho.ncg, dtype=bool)
    for pp in range(ho.np):
        v = ho.ghost_view(im.view(np.uint8), pp)
        v[(slice(1, -1),) * ndim] = 1
    for step in range(2):
        ra, sa = a.advance(1)
        rb, sb = b.advance(1)
        assert ra == rb == 0, (step, sa, sb)
        assert sa["t"] == sb["t"] and sa["nst"] == sb["nst"], (step, sa, sb)
        ya, yb = a.get_state(), b.get_state()
        for m in range(a.nc):
            e = _rel(yb[m * ho.ncg:(m + 1) * ho.ncg][im], ya[m * ho.ncg:(m + 1) * ho.ncg][im])
            assert e < 1e-8, (step, m, e)
    print(case, "fp64 nli", sa["nli"], "fp32 nli", sb["nli"])
