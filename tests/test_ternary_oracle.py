"""Pins of the oracle's three-phase ternary multi-order restatement
(oracle/oracle_ternary.c) that do not retype its formulas:

- the closed-form phase concentrations satisfy the two KKS conditions they
  solve: one diffusion potential 2 A_ak (x_ak - e_ak) in the three phases and
  the mass balance c_k = sum_a h_a x_ak (random wells, weights and c);
- the driving force is the derivative of the grand-potential density
  sum_a h_a(phi) w_a, w_a = f_a - mu_a . c_a (fixed), with respect to each
  order parameter, h_a = (sum of a's phi^2) / sum phi^2 (central finite
  differences): a wrong sign, weight or phase assignment fails it;
- the three-phase composition RHS conserves each species on a periodic grid
  and, with one phase everywhere (uniform diffusivity), is M_c D times the
  5-point Laplacian of that phase's concentration (a quadratic field gives
  the constant 2 M_c D (a + b));
- the multi-spheres generator gives the reference's Sintering initial state:
  four discs of 448 cells (0.109375 of the 64^2 grid) with the liquid order
  parameter 1 - the others;
- the Sintering regression (tests/Sintering/test2d.py): at t = 0.006 s the
  largest and smallest order-parameter volume fractions are 0.586 and 0.118
  (+- 0.001), reached by the oracle integrator converged (Newton atol 1e-6,
  the deck's 1e-4 leaves a Newton-tolerance drift of the composition
  integral of ~3e-3 with this preconditioner).
"""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P
from oracle import oracle as O


def _tern_params(rng):
    p, _, _ = P.sintering_deck()
    for a in range(3):
        for k in range(2):
            p.tern_A[a][k] = float(rng.uniform(1e3, 1e4))
            p.tern_ceq[a][k] = float(rng.uniform(0.05, 0.5))
    p.vm_L, p.vm_A, p.vm_B = 1.5e-5, 1.2e-5, 1.8e-5
    return p


def test_phase_concentrations_satisfy_kks_conditions():
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = _tern_params(rng)
        w = rng.uniform(0.01, 1.0, 3)
        h = w / w.sum()
        c = rng.uniform(0.0, 0.6, 2)
        x = np.zeros(6)
        O.lib().o_tern3_phase_conc(C.byref(p), O.ptr(np.ascontiguousarray(c)),
                                   O.ptr(np.ascontiguousarray(h)), O.ptr(x))
        for k in range(2):
            mu = [2 * p.tern_A[a][k] * (x[2 * a + k] - p.tern_ceq[a][k]) for a in range(3)]
            assert abs(mu[0] - mu[1]) < 1e-9 * max(1.0, abs(mu[0]))
            assert abs(mu[0] - mu[2]) < 1e-9 * max(1.0, abs(mu[0]))
            assert abs(sum(h[a] * x[2 * a + k] for a in range(3)) - c[k]) < 1e-13


def _grand(p, phi, ph, norder):
    """sum_a h_a(phi) w_a with w from the phase concentrations ph."""
    nA = p.norderp_A
    s = (phi ** 2).sum()
    h = [phi[norder - 1] ** 2 / s, (phi[:nA] ** 2).sum() / s, (phi[nA:norder - 1] ** 2).sum() / s]
    vm = [p.vm_L, p.vm_A, p.vm_B]
    g = 0.0
    for a in range(3):
        fac = 1e-6 / vm[a]
        xa = ph[2 * a:2 * a + 2]
        f = sum(p.tern_A[a][k] * (xa[k] - p.tern_ceq[a][k]) ** 2 for k in range(2)) * fac
        mu = [2 * p.tern_A[a][k] * (xa[k] - p.tern_ceq[a][k]) * fac for k in range(2)]
        g += h[a] * (f - mu[0] * xa[0] - mu[1] * xa[1])
    return g


@pytest.mark.parametrize("norder,nA", [(5, 2), (3, 1), (4, 1)])
def test_driving_force_is_grand_potential_derivative(norder, nA):
    rng = np.random.default_rng(11 + norder)
    for _ in range(30):
        p = _tern_params(rng)
        p.norderp, p.norderp_A = norder, nA
        phi = rng.uniform(0.05, 1.0, norder)
        ph = rng.uniform(0.0, 0.6, 6)
        for ip in range(norder):
            got = O.lib().o_tern3_driving_force(C.byref(p), norder, O.ptr(phi), O.ptr(ph), ip)
            e = 1e-6
            a, b = phi.copy(), phi.copy()
            a[ip] += e
            b[ip] -= e
            fd = (_grand(p, a, ph, norder) - _grand(p, b, ph, norder)) / (2 * e)
            assert abs(got - fd) < 1e-6 * max(1.0, abs(fd)), (ip, got, fd)


def test_composition_rhs_conserves_and_reduces_to_laplacian():
    n = 24
    p, _, _ = P.sintering_deck()
    p.D_liquid = 3.0
    p.conc_mobility = 1.7
    pt = O.Patch(2, (n, n, 1), g=1, dx=(0.05, 0.05, 1.0))
    rng = np.random.default_rng(5)
    G = pt.G
    # random phases: conservation on the periodic grid
    phi = pt.cells(5)
    pt.interior(phi)[...] = rng.uniform(0, 1, (5, 1, n, n))
    ph = pt.cells(6)
    pt.interior(ph)[...] = rng.uniform(0, 0.5, (6, 1, n, n))
    O.lib().o_ghost_fill(C.byref(G), O.ptr(phi), 5)
    O.lib().o_ghost_fill(C.byref(G), O.ptr(ph), 6)
    DL, DA, DB = (np.zeros(pt.nf) for _ in range(3))
    O.lib().o_ebs3_coeffs(C.byref(G), C.byref(p), 5, O.ptr(phi), O.ptr(DL), O.ptr(DA), O.ptr(DB))
    rhs = pt.cells(2)
    O.lib().o_conc_rhs_ebs3(C.byref(G), C.byref(p), O.ptr(ph), O.ptr(DL), O.ptr(DA), O.ptr(DB),
                            O.ptr(rhs))
    for k in range(2):
        r = pt.interior(rhs[k])
        assert abs(r.sum()) < 1e-10 * np.abs(r).sum()
    # the liquid everywhere: D_L uniform, rhs = M_c D_L lap(cL_k)
    phi[:] = 0.0
    phi[4] = 1.0
    y, x = np.mgrid[0:n, 0:n] * 0.05
    a, b = 0.3, -0.7
    for k in range(2):
        pt.interior(ph[k])[...] = (k + 1) * (a * x * x + b * y * y)
    O.lib().o_ebs3_coeffs(C.byref(G), C.byref(p), 5, O.ptr(phi), O.ptr(DL), O.ptr(DA), O.ptr(DB))
    assert np.allclose(DL, 3.0) and np.all(DA == 0) and np.all(DB == 0)
    O.lib().o_conc_rhs_ebs3(C.byref(G), C.byref(p), O.ptr(ph), O.ptr(DL), O.ptr(DA), O.ptr(DB),
                            O.ptr(rhs))
    for k in range(2):
        r = pt.interior(rhs[k])[0, 1:-1, 1:-1]
        want = 1.7 * 3.0 * 2.0 * (k + 1) * (a + b)
        assert np.max(np.abs(r - want)) < 1e-8 * abs(want)


def test_multi_spheres_initial_state():
    p, (phi, c), chk = P.sintering_deck()
    assert phi.shape == (5, 64, 64) and c.shape == (2, 64, 64)
    for g in range(4):
        assert int(phi[g].sum()) == 448
    assert np.all(phi.sum(axis=0) == 1.0)
    # brute force of one disc: cell centres (i + 0.5, j + 0.5, 0.5), centre z = 1
    j, i = np.mgrid[0:64, 0:64]
    inside = (i + 0.5 - 44) ** 2 + (j + 0.5 - 20) ** 2 + 0.25 < 144
    assert np.array_equal(phi[1] == 1.0, inside)
    assert np.array_equal(c[0], np.where(phi[0] + phi[1] > 0, 1.0, 0.0))
    assert np.array_equal(c[1], np.where(phi[2] + phi[3] > 0, 1.0, 0.0))


def sintering_run(sim, get_phi, end_time):
    """Advance until t > end_time; returns (volume fractions, t, steps)."""
    t, n = 0.0, 0
    while t <= end_time and n < 2000:
        rc, st = sim.advance(1)
        assert rc == 0, st
        t, n = st["t"], n + 1
    phi = get_phi()
    return phi.reshape(phi.shape[0], -1).mean(axis=1), t, n


def test_sintering_regression_oracle():
    p, (phi, c), chk = P.sintering_deck()
    s = P.solver_defaults(dt=5e-4, dt_init=1e-6, adaptive=1, atol=1e-6, rtol=1e-8, maxl=10,
                          max_newton=3)
    sim = O.Sim(p, s, 64, True)
    sim.set_state(phi, c, None)
    vol, t, n = sintering_run(sim, lambda: sim.get_state()[0], chk["end_time"])
    assert abs(vol.max() - chk["vmax"]) <= chk["tol"], vol
    assert abs(vol.min() - chk["vmin"]) <= chk["tol"], vol
    # the composition integral is kept
    cc = sim.get_state()[1]
    assert abs(cc.mean() - c.mean()) < 1e-4
