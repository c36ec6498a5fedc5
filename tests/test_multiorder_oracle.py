"""CPU checks of the oracle's multi-order phase right-hand side
(MultiOrderRHSStrategy, computerhsmultiorder) and of projectphi, pinned by
hand-computed special cases: a single grain is a steady state, the uniform
50/50 mixture of two orders has g' = 0.5^3 - 0.5 + 2 (0.5)(1.5)(0.5 - 0.5) =
-0.375 (with the reference's sum - phi; sum - phi^2 would give 0), the
diffusion term is the 5-point Laplacian of eps^2 / 2 (symbol (2 cos kh - 2)/h^2
on a cosine), and the projection lands on sum(phi) = 1 with err of zero mean."""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P
from oracle import oracle as O


def _rhs(pt, p, phi, M=None):
    n = phi.shape[0]
    r = pt.cells(n)
    O.lib().o_phase_rhs_multiorder(C.byref(pt.G), C.byref(p), n, O.ptr(phi),
                                   O.ptr(M) if M is not None else None, O.ptr(r))
    return r


@pytest.mark.parametrize("ndim", [2, 3])
def test_uniform_special_cases(ndim):
    p = P.two_grain_params(ndim=ndim)
    pt = O.Patch(ndim, (8, 6, 4))
    m = 8.0 * p.phi_well_scale
    phi = pt.cells(3)
    phi[0] = 1.0
    assert np.all(pt.interior(_rhs(pt, p, phi)) == 0.0)
    phi = pt.cells(2)
    phi[:] = 0.5
    r = pt.interior(_rhs(pt, p, phi))
    assert np.allclose(r, 0.375 * m * p.phi_mobility, rtol=1e-15, atol=0)


@pytest.mark.parametrize("ndim", [2, 3])
def test_diffusion_term_symbol(ndim):
    """w = 0: rhs_1 = M (eps^2 / 2) sum_d (2 cos(k_d h_d) - 2) / h_d^2 cos(k_d x_d)
    for phi_1 = sum_d cos(k_d x_d); a uniform second order gets rhs 0."""
    dx = (0.5, 0.25, 0.125)
    p = P.two_grain_params(ndim=ndim, dx=dx)
    p.phi_well_scale = 0.0
    pt = O.Patch(ndim, (12, 10, 8), dx=dx)
    k3, j3, i3 = np.meshgrid(*[np.arange(s) for s in pt.gshape], indexing="ij")
    x = ((i3 - pt.g + 0.5) * dx[0], (j3 - pt.g + 0.5) * dx[1], (k3 - pt.gz + 0.5) * dx[2])
    n = (pt.nx, pt.ny, pt.nz)
    phi = pt.cells(2)
    want = np.zeros(pt.gshape)
    for d in range(ndim):
        k = 2 * np.pi * (d + 1) / (n[d] * dx[d])
        phi[0] += np.cos(k * x[d])
        want += (2 * np.cos(k * dx[d]) - 2) / dx[d] ** 2 * np.cos(k * x[d])
    phi[1] = 0.3
    want *= p.phi_mobility * p.epsilon_phi ** 2 / 2
    r = _rhs(pt, p, phi)
    err = np.abs(pt.interior(r[0]) - pt.interior(want)).max() / np.abs(pt.interior(want)).max()
    assert err < 1e-12
    assert np.all(pt.interior(r[1]) == 0.0)
    # a mobility field multiplies cell by cell
    M = np.full(pt.gshape, 2.0)[None].copy()
    r2 = _rhs(pt, p, phi, M)
    assert np.allclose(pt.interior(r2[0]), 2.0 / p.phi_mobility * pt.interior(r[0]),
                       rtol=1e-14, atol=1e-14)


def test_projection():
    pt = O.Patch(2, (8, 6, 1))
    rng = np.random.default_rng(3)
    phi = rng.uniform(0, 1, (3,) + pt.gshape)
    err = rng.normal(size=(3,) + pt.gshape)
    err0 = err.copy()
    corr = pt.cells(3)
    O.lib().o_project_phases(C.byref(pt.G), 3, O.ptr(phi), O.ptr(corr), O.ptr(err))
    s = pt.interior(phi + corr).sum(0)
    assert np.abs(s - 1.0).max() < 1e-14
    assert np.abs(pt.interior(err).sum(0)).max() < 1e-14
    d = pt.interior(err0 - err)
    assert np.abs(d - d[:1]).max() < 1e-14
