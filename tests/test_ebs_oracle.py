"""CPU checks of the EBS composition right-hand side in the oracle
(EBSCompositionRHSStrategy: flux = sum over phases of D_phase grad c_phase,
with D_phase carrying the phase weight).  Pinned by closed forms: the discrete
divergence of the flux of a quadratic phase composition, conservation on a
periodic patch, and the reduction to the KKS-form right-hand side when both
phases share the composition and phi is uniform (then D_L + D_A is the KKS
mixture diffusivity and the phi-gradient flux vanishes)."""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P
from oracle import oracle as O


def _ebs(pt, p, phi, cl, cs):
    DL, DA = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_ebs_coeffs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(DL), O.ptr(DA))
    r = pt.cells()
    O.lib().o_conc_rhs_ebs(C.byref(pt.G), C.byref(p), O.ptr(cl), O.ptr(cs), O.ptr(DL),
                           O.ptr(DA), O.ptr(r))
    return DL, DA, r


def _diffusivities(p):
    dl = p.D_liquid * np.exp(-p.Q0_liquid / (P.R_GAS * p.temperature))
    ds = p.D_solid * np.exp(-p.Q0_solid / (P.R_GAS * p.temperature))
    return dl, ds


def _centres(pt, dx):
    """cell-centre coordinates over the whole ghosted array."""
    gz = pt.gz
    k, j, i = np.meshgrid(*[np.arange(s) for s in pt.gshape], indexing="ij")
    return ((i - pt.g + 0.5) * dx[0], (j - pt.g + 0.5) * dx[1], (k - gz + 0.5) * dx[2])


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("phase", ["solid", "liquid"])
def test_single_phase_quadratic(ndim, phase):
    """phi = 1 (solid) or 0 (liquid) everywhere: only that phase's diffusivity
    survives, and for c_phase = a x^2 + b y^2 (+ e z^2) the discrete flux
    divergence is 2 (a + b (+ e)) exactly, so rhs = M D_phase 2 (a + b + e)."""
    dx = (0.5, 0.25, 0.125)
    p = P.two_grain_params(ndim=ndim, dx=dx)
    pt = O.Patch(ndim, (12, 10, 8), dx=dx)
    x, y, z = _centres(pt, dx)
    a, b, e = 0.3, -0.7, 1.1
    quad = a * x ** 2 + b * y ** 2 + (e * z ** 2 if ndim == 3 else 0.0)
    lin = 0.2 + 0.1 * x - 0.3 * y
    phival = 1.0 if phase == "solid" else 0.0
    phi = np.full(pt.gshape, phival)[None]
    cl = (lin if phase == "solid" else quad)[None].copy()
    cs = (quad if phase == "solid" else lin)[None].copy()
    DL, DA, r = _ebs(pt, p, phi, cl, cs)
    dl, ds = _diffusivities(p)
    if phase == "solid":
        assert np.all(DL == 0.0) and np.allclose(DA, ds, rtol=1e-15, atol=0)
        d = ds
    else:
        assert np.all(DA == 0.0) and np.allclose(DL, dl, rtol=1e-15, atol=0)
        d = dl
    want = p.conc_mobility * d * 2.0 * (a + b + (e if ndim == 3 else 0.0))
    ri = pt.interior(r)
    assert np.max(np.abs(ri - want)) < 1e-9 * abs(want)


@pytest.mark.parametrize("ndim", [2, 3])
def test_conservation_periodic(ndim):
    rng = np.random.Generator(np.random.Philox(31))
    pt = O.Patch(ndim, (16, 12, 10))
    p = P.two_grain_params(ndim=ndim)
    shape = (pt.nz, pt.ny, pt.nx)
    phi = pt.load(pt.cells(), rng.uniform(0, 1, shape))
    cl = pt.load(pt.cells(), rng.uniform(0.2, 0.6, shape))
    cs = pt.load(pt.cells(), rng.uniform(0.2, 0.6, shape))
    DL, DA, r = _ebs(pt, p, phi, cl, cs)
    ri = pt.interior(r)
    assert abs(ri.sum()) < 1e-12 * np.abs(ri).sum()
    assert np.abs(ri).max() > 0
    dl, ds = _diffusivities(p)
    # linear weights: D_L / dl + D_A / ds = 1 face by face
    assert np.max(np.abs(DL / dl + DA / ds - 1.0)) < 1e-14


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("interp", ["l", "p"])
def test_reduces_to_kks_form(ndim, interp):
    """cl = cs = c and uniform phi: EBS rhs = KKS-form rhs when the diffusion
    and energy interpolations agree."""
    rng = np.random.Generator(np.random.Philox(37))
    pt = O.Patch(ndim, (14, 12, 8))
    p = P.two_grain_params(ndim=ndim)
    p.energy_interp = interp.encode()
    p.diff_interp = interp.encode()
    shape = (pt.nz, pt.ny, pt.nx)
    phi = pt.load(pt.cells(), np.full(shape, 0.37))
    c = pt.load(pt.cells(), rng.uniform(0.2, 0.6, shape))
    DL, DA, r = _ebs(pt, p, phi, c, c)
    Dc, Dp = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_conc_coeffs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(c), O.ptr(c),
                          O.ptr(Dc), O.ptr(Dp))
    rk = pt.cells()
    O.lib().o_conc_rhs(C.byref(pt.G), C.byref(p), O.ptr(c), O.ptr(phi), O.ptr(Dc), O.ptr(Dp),
                       O.ptr(rk))
    assert np.max(np.abs(DL + DA - Dc)) < 1e-15 * np.abs(Dc).max()
    a, b = pt.interior(r), pt.interior(rk)
    assert np.linalg.norm(a - b) < 1e-12 * np.linalg.norm(b)


def test_diffusion_interpolation_is_its_own_parameter():
    """h of the diffusivities follows diff_interp, not energy_interp."""
    pt = O.Patch(2, (8, 8, 1))
    p = P.two_grain_params(ndim=2)
    phi = pt.load(pt.cells(), np.full((1, 8, 8), 0.25))
    dl, ds = _diffusivities(p)
    for ch, h in ((b"l", 0.25), (b"p", 0.25 ** 3 * (10 - 15 * 0.25 + 6 * 0.25 ** 2))):
        p.diff_interp = ch
        DL, DA, _ = _ebs(pt, p, phi, phi, phi)
        assert np.allclose(DA, h * ds, rtol=1e-14, atol=0)
        assert np.allclose(DL, (1 - h) * dl, rtol=1e-14, atol=0)


def test_integrator_ebs_conserves_composition():
    """A few oracle integrator steps with the EBS form on a periodic 2-D
    two-grain state keep the mean composition."""
    n = 32
    p = P.two_grain_params(ndim=2, quad_ceq=(0.40, 0.45), conc_rhs_form=P.CONC_RHS_EBS)
    phi, c, q = P.two_grain_state(n=n, radius=10.0)
    s = P.solver_defaults(dt=1e-3, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8)
    sim = O.Sim(p, s, n)
    sim.set_state(phi, c, q)
    rc, st = sim.advance(3)
    assert rc == 0
    _, c1, _, cl, cs = sim.get_state()
    assert abs(c1.mean() - c.mean()) < 1e-12
    assert np.abs(c1 - c).max() > 0


def _ebs_iso(pt, p, phi, cl, cs):
    r = pt.cells()
    O.lib().o_conc_rhs_ebs_iso(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(cl), O.ptr(cs),
                               O.ptr(r))
    return r


@pytest.mark.parametrize("ndim", [2, 3])
def test_isotropic_stencil_reduces_to_regular_on_one_dimensional_fields(ndim):
    """add_flux_iso's weights sum to one (10/12 + 2/12 in 2-D, 8/12 + 4/12 in
    3-D): on fields varying along x only every transverse face carries the
    same diffusivity and difference, so the isotropic and regular stencils
    agree to round-off; with variation along y as well they differ."""
    dx = (0.5, 0.25, 0.125)
    p = P.two_grain_params(ndim=ndim, dx=dx, conc_rhs_form=P.CONC_RHS_EBS, diff_interp=b"p")
    pt = O.Patch(ndim, (12, 10, 8), dx=dx)
    rng = np.random.default_rng(3)
    prof = lambda: rng.uniform(0.1, 0.9, pt.gshape[-1])[None, None, :]
    phi = np.broadcast_to(prof(), pt.gshape).copy()[None]
    cl = np.broadcast_to(prof(), pt.gshape).copy()[None]
    cs = np.broadcast_to(prof(), pt.gshape).copy()[None]
    for a in (phi, cl, cs):
        O.lib().o_ghost_fill_periodic(C.byref(pt.G), O.ptr(a), 1)
    r_reg = _ebs(pt, p, phi, cl, cs)[2]
    r_iso = _ebs_iso(pt, p, phi, cl, cs)
    err = np.abs(pt.interior(r_iso - r_reg)).max() / np.abs(pt.interior(r_reg)).max()
    assert err < 1e-13
    cl2 = cl + 0.05 * rng.normal(size=cl.shape)
    O.lib().o_ghost_fill_periodic(C.byref(pt.G), O.ptr(cl2), 1)
    assert np.abs(pt.interior(_ebs_iso(pt, p, phi, cl2, cs) - _ebs(pt, p, phi, cl2, cs)[2])).max() > 1e-8


@pytest.mark.parametrize("ndim", [2, 3])
def test_isotropic_stencil_conserves(ndim):
    """Periodic patch: the flux divergence sums to zero."""
    dx = (0.5, 0.25, 0.125)
    p = P.two_grain_params(ndim=ndim, dx=dx, conc_rhs_form=P.CONC_RHS_EBS, diff_interp=b"p")
    pt = O.Patch(ndim, (12, 10, 8), dx=dx)
    rng = np.random.default_rng(5)
    phi, cl, cs = (pt.load(pt.cells(), rng.uniform(0, 1, (pt.nz, pt.ny, pt.nx))) for _ in range(3))
    r = _ebs_iso(pt, p, phi, cl, cs)
    assert abs(pt.interior(r).sum()) < 1e-12 * np.abs(pt.interior(r)).sum()


def _ebs_4th(pt, p, phi, cl, cs):
    DL, DA = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_ebs_coeffs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(DL), O.ptr(DA))
    r = pt.cells()
    O.lib().o_conc_rhs_ebs_4th(C.byref(pt.G), C.byref(p), O.ptr(cl), O.ptr(cs), O.ptr(DL),
                               O.ptr(DA), O.ptr(r))
    return r


@pytest.mark.parametrize("ndim", [2, 3])
def test_order4_stencil_fourier_symbol(ndim):
    """add_flux_4th on a periodic patch: the divergence of the face fluxes is
    the fourth-order Laplacian (-1, 16, -30, 16, -1) / 12 dx^2, whose symbol on
    cos(k x) is (32 cos(k dx) - 2 cos(2 k dx) - 30) / (12 dx^2).  phi = 0
    leaves the liquid's uniform diffusivity only, so for cL = sum_d cos(k_d x_d)
    rhs = M D_L sum_d symbol_d cos(k_d x_d) (cS, with D_A = 0, drops out)."""
    dx = (0.5, 0.25, 0.125)
    p = P.two_grain_params(ndim=ndim, dx=dx, conc_rhs_form=P.CONC_RHS_EBS)
    p.ebs_stencil = P.EBS_ORDER4
    pt = O.Patch(ndim, (12, 10, 8), g=2, dx=dx)
    x = _centres(pt, dx)
    n = (pt.nx, pt.ny, pt.nz)
    rng = np.random.default_rng(11)
    cl = np.zeros(pt.gshape)
    expect = np.zeros(pt.gshape)
    for d in range(ndim):
        k = 2 * np.pi * (d + 1) / (n[d] * dx[d])
        cl += np.cos(k * x[d])
        sym = (32 * np.cos(k * dx[d]) - 2 * np.cos(2 * k * dx[d]) - 30) / (12 * dx[d] ** 2)
        expect += sym * np.cos(k * x[d])
    phi = np.zeros((1,) + pt.gshape)
    cs = rng.uniform(0, 1, (1,) + pt.gshape)
    r = _ebs_4th(pt, p, phi, cl[None].copy(), cs)
    dl, _ = _diffusivities(p)
    want = p.conc_mobility * dl * expect
    err = np.abs(pt.interior(r[0]) - pt.interior(want)).max() / np.abs(pt.interior(want)).max()
    assert err < 1e-12


@pytest.mark.parametrize("ndim", [2, 3])
def test_order4_physical_boundary(ndim):
    """Non-periodic patch with zero-slope ghosts: the boundary faces keep the
    second-order difference (zero there), so the divergence sums to zero; a
    fourth-order difference across those faces would not (it reads the
    second cell inside).  A quartic profile in x gets its exact Laplacian
    12 a x^2 in the cells whose faces are all interior (the stencil is exact
    to degree five)."""
    dx = (0.5, 0.25, 0.125)
    p = P.two_grain_params(ndim=ndim, dx=dx, conc_rhs_form=P.CONC_RHS_EBS)
    p.ebs_stencil = P.EBS_ORDER4
    pt = O.Patch(ndim, (12, 10, 8), g=2, dx=dx)
    pt.G.periodic = 0
    rng = np.random.default_rng(13)
    phi, cl, cs = (rng.uniform(0, 1, (1,) + pt.gshape) for _ in range(3))
    for a in (phi, cl, cs):
        O.lib().o_ghost_fill(C.byref(pt.G), O.ptr(a), 1)
    r = _ebs_4th(pt, p, phi, cl, cs)
    ri = pt.interior(r)
    assert abs(ri.sum()) < 1e-12 * np.abs(ri).sum()
    # quartic in x, liquid only
    x = _centres(pt, dx)[0]
    a = 0.7
    phi0 = np.zeros((1,) + pt.gshape)
    q = (a * (x - 2.0) ** 4)[None].copy()
    O.lib().o_ghost_fill(C.byref(pt.G), O.ptr(q), 1)
    r = _ebs_4th(pt, p, phi0, q, q.copy())
    dl, _ = _diffusivities(p)
    want = p.conc_mobility * dl * 12 * a * (x - 2.0) ** 2
    ri, wi = pt.interior(r[0]), pt.interior(want)
    inner = ri[..., 2:-2], wi[..., 2:-2]
    assert np.abs(inner[0] - inner[1]).max() < 1e-9 * np.abs(inner[1]).max()
    assert np.abs(ri[..., :2] - wi[..., :2]).max() > 1e-6


def test_integrator_order4_conserves_composition():
    """Oracle integrator with the order4 stencil (two ghost cells) on a
    non-periodic 2-D two-grain state keeps the mean composition."""
    n = 32
    p = P.two_grain_params(ndim=2, quad_ceq=(0.40, 0.45), conc_rhs_form=P.CONC_RHS_EBS)
    p.ebs_stencil = P.EBS_ORDER4
    phi, c, q = P.two_grain_state(n=n, radius=10.0)
    s = P.solver_defaults(dt=1e-3, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8)
    sim = O.Sim(p, s, n, periodic=False)
    sim.set_state(phi, c, q)
    rc, st = sim.advance(3)
    assert rc == 0
    _, c1, _, cl, cs = sim.get_state()
    assert abs(c1.mean() - c.mean()) < 1e-12
    assert np.abs(c1 - c).max() > 0
