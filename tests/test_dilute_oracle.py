"""CPU checks of the dilute binary solution model (KKSdiluteBinary) in the
oracle: the reference's KKS and equilibrium unit-test cases
(tests/testKKSdiluteBinaryKKS.cc, tests/testKKSdiluteBinaryEquilibrium.cc with
their test.input values) and closed forms of the model."""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P
from oracle import oracle as O

TOL_REF = 1.0e-5  # both reference tests compare dfL/dc and dfS/dc to 1e-5


def _mu(p, c, ph):
    return O.lib().o_dfenergy(C.byref(p), c, ph)


def test_reference_kks_case():
    """testKKSdiluteBinaryKKS: T = 1710 K, c = 0.05, phi = 0.5, guess (0.5, 0.5)."""
    p = P.dilute_params(ndim=2)
    xl, xs = np.array([0.5]), np.array([0.5])
    it = O.lib().o_kks_cell(C.byref(p), 0.05, 0.5, O.ptr(xl), O.ptr(xs))
    assert it >= 0
    assert abs(_mu(p, xl[0], 0) - _mu(p, xs[0], 1)) < TOL_REF
    h = 0.5  # linear interpolation
    assert abs((1 - h) * xl[0] + h * xs[0] - 0.05) < 1e-15
    assert abs(xs[0] / xl[0] - 0.7965) < 1e-14


def test_reference_equilibrium_case():
    """testKKSdiluteBinaryEquilibrium at 1710 K: the two-phase equilibrium lies
    on the liquidus c_L = (T - T_m)/m, c_S = k_e c_L, with equal chemical
    potentials and equal grand potentials (common tangent)."""
    p = P.dilute_params(ndim=2)
    ceL = (p.temperature - p.dilute_Tm) / p.dilute_m
    ceS = p.dilute_keq * ceL
    muL, muS = _mu(p, ceL, 0), _mu(p, ceS, 1)
    assert abs(muL - muS) < TOL_REF
    fL = O.lib().o_fenergy(C.byref(p), ceL, 0)
    fS = O.lib().o_fenergy(C.byref(p), ceS, 1)
    gL, gS = fL - muL * ceL, fS - muS * ceS
    assert abs(gL - gS) < 1e-9 * abs(gL)


@pytest.mark.parametrize("ph", [0, 1])
def test_derivatives_finite_differences(ph):
    p = P.dilute_params(ndim=2)
    for c in (0.01, 0.05, 0.2, 0.6):
        e = 1e-6 * c
        f = lambda x: O.lib().o_fenergy(C.byref(p), x, ph)  # noqa: E731
        d = lambda x: O.lib().o_dfenergy(C.byref(p), x, ph)  # noqa: E731
        fd1 = (f(c + e) - f(c - e)) / (2 * e)
        fd2 = (d(c + e) - d(c - e)) / (2 * e)
        assert abs(fd1 - d(c)) < 1e-6 * max(1.0, abs(d(c)))
        assert abs(fd2 - O.lib().o_d2fenergy(C.byref(p), c, ph)) < 1e-6 * abs(fd2)


def test_driving_force_vanishes_at_equilibrium():
    """A flat two-phase state at the equilibrium compositions has no KKS
    driving force: (f_L - f_S) - mu_S (c_L - c_S) = 0 (the reference's
    addDrivingForce with mu = mu_A(c_A))."""
    p = P.dilute_params(ndim=2)
    ceL = (p.temperature - p.dilute_Tm) / p.dilute_m
    ceS = p.dilute_keq * ceL
    fL = O.lib().o_fenergy(C.byref(p), ceL, 0)
    fS = O.lib().o_fenergy(C.byref(p), ceS, 1)
    df = (fL - fS) - _mu(p, ceS, 1) * (ceL - ceS)
    assert abs(df) < 1e-9 * abs(fL)


def test_kim_mobility_closed_forms():
    """Kim's KKS mobility (source/KimMobilityStrategy*.cc) for the dilute model:
    f_L'' = RT / c, so with c_S = k c_L the infinite-interface-mobility value is
    factor / ((1-k)^2 c_L RT / D_L * 1e-6 / v_m); the finite value is
    1 / alpha; the anti-trapping value is their series combination
    1/M = 1/M_finite + 1/M_inf (the two resistances add)."""
    base = P.dilute_params(ndim=3, phi_mobility_type=P.PHI_MOB_KIM, D_liquid=1.0e3,
                           Q0_liquid=0.0, energy_interp=b"p")
    R, T = P.R_GAS, base.temperature
    eps, w = base.epsilon_phi, base.phi_well_scale
    factor = 16.0 * w / (3.0 * eps * eps * (47.0 / 60.0))
    k = base.dilute_keq
    for cl in (0.02, 0.05, 0.1):
        cs = k * cl
        zeta = (1 - k) ** 2 * cl * R * T / base.D_liquid * 1e-6 / base.vm_L
        m_inf = O.lib().o_phase_mobility(C.byref(base), cl, cs)
        assert abs(m_inf - factor / zeta) < 1e-12 * m_inf
        mu = P.kim_interface_mobility(base.vm_L, base.dilute_Tm, base.dilute_m, k, 2.0e-5)
        fin = P.dilute_params(ndim=3, phi_mobility_type=P.PHI_MOB_KIM, D_liquid=1.0e3,
                              Q0_liquid=0.0, energy_interp=b"p", kim_interface_mobility=mu)
        m_fin = O.lib().o_phase_mobility(C.byref(fin), cl, cs)
        alpha = 3.0 * np.sqrt(2.0) * (eps / np.sqrt(16.0 * w)) / mu
        assert abs(m_fin - 1.0 / alpha) < 1e-12 * m_fin
        fin.kim_antitrapping = 1
        m_at = O.lib().o_phase_mobility(C.byref(fin), cl, cs)
        assert abs(1.0 / m_at - (1.0 / m_fin + 1.0 / m_inf)) < 1e-10 / m_at
    const = P.dilute_params(ndim=3)
    assert O.lib().o_phase_mobility(C.byref(const), 0.05, 0.04) == const.phi_mobility
