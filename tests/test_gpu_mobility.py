"""Composition-dependent EBS diffusivities (CALPHAD atomic mobilities times
d2f/dc2, MobilityCompositionDiffusionStrategy) on the device against the
oracle: the face coefficients on 2-D and 3-D levels (wider than one tile),
the integrator on the reference's SingleGrainGrowthAuNi deck step by step,
and that deck's regression values (solid fraction 0.32 +- 0.01 after
t = 0.3 s, composition integral within 1e-4) on the device."""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu

N = 64


def _dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _p(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("ndim,shape", [(2, (40, 24, 1)), (3, (20, 12, 10)), (3, (70, 12, 6))])
def test_mobility_coeffs_match_oracle(dev, ndim, shape):
    from oracle import oracle as O
    from ampe_amd import hip as H
    p, _, _, _ = P.single_grain_growth_auni_deck()
    p.ndim = ndim
    p.temperature = 1320.0
    p.calphad = P.auni_phases(1320.0)
    nx, ny, nz = shape
    pt = O.Patch(ndim, shape, dx=tuple(p.dx))
    box = P.Box()
    box.lo[:] = [0, 0, 0]
    box.hi[:] = [nx - 1, ny - 1, nz - 1 if ndim == 3 else 0]
    lev = H.Level(ndim, [box], 1, periodic=box)
    rng = np.random.Generator(np.random.Philox(41))
    sh = (pt.nz, pt.ny, pt.nx)
    phi = pt.load(pt.cells(), rng.uniform(0.0, 1.0, sh))
    cl = pt.load(pt.cells(), rng.uniform(0.05, 0.6, sh))
    cs = pt.load(pt.cells(), rng.uniform(0.05, 0.6, sh))
    DL_o, DA_o = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_ebs_coeffs_mob(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(cl), O.ptr(cs),
                             O.ptr(DL_o), O.ptr(DA_o))
    phd, cld, csd = (_dev(a, dev) for a in (phi, cl, cs))
    DL, DA = _dev(np.zeros(pt.nf), dev), _dev(np.zeros(pt.nf), dev)
    H.check(H.lib().ampe_conc_coeffs_ebs_mobility(lev.h, C.byref(p), _p(phd), _p(cld), _p(csd),
                                                  _p(DL), _p(DA), None))
    for got, want in ((DL.cpu().numpy(), DL_o), (DA.cpu().numpy(), DA_o)):
        assert np.isfinite(want).all() and np.abs(want).max() > 0
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


def _deck_sims(s):
    from oracle import oracle as O
    from ampe_amd import hip as H
    p, ramp, st, ck = P.single_grain_growth_auni_deck()
    box = H.boxes_of((N, N), 2)
    g = H.Sim(p, s, [box], periodic=None)
    g.set_temperature_ramp(ramp["T0"], ramp["dTdt"], ramp["T_target"], ramp["tab"])
    g.set_state(*st)
    o = O.Sim(p, s, (N, N, 1), periodic=False)
    o.set_temperature_ramp(ramp["T0"], ramp["dTdt"], ramp["T_target"], ramp["tab"])
    o.set_state(*st)
    return p, ck, g, o


def test_grain_growth_deck_matches_oracle(dev):
    s = P.kks_composition_solver(dt=2e-2, adaptive=1, atol=1e-6, rtol=1e-6)
    p, ck, g, o = _deck_sims(s)
    for step in range(6):
        rg, sg = g.advance(5)
        ro, so = o.advance(5)
        assert rg == 0 and ro == 0
        assert abs(sg["t"] - so["t"]) <= 1e-9 * so["t"]
        assert (sg["nst"], sg["nni"]) == (so["nst"], so["nni"]), (step, sg, so)
        fg = g.get_state((N, N))
        fo = o.get_state()
        for name, a, b in zip(("phi", "c", "q", "cl", "cs"), fg, fo):
            if name == "q":
                continue
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert err < 1e-9, (step, name, err)
    print("t", sg["t"], "nst", sg["nst"], "nli", sg["nli"])


def test_single_grain_growth_auni_regression(dev):
    """The device run of tests/SingleGrainGrowthAuNi to t >= 0.3 s with the
    adaptive BDF at atol 1e-6 (the oracle run of the same settings reaches
    solid fraction 0.321, composition drift 3.6e-5)."""
    s = P.kks_composition_solver(dt=2e-2, adaptive=1, atol=1e-6, rtol=1e-6)
    from ampe_amd import hip as H
    p, ramp, st, ck = P.single_grain_growth_auni_deck()
    box = H.boxes_of((N, N), 2)
    g = H.Sim(p, s, [box], periodic=None)
    g.set_temperature_ramp(ramp["T0"], ramp["dTdt"], ramp["T_target"], ramp["tab"])
    g.set_state(*st)
    dx = p.dx[0]
    I0 = g.get_state((N, N))[1].sum() * dx * dx
    t = 0.0
    while t < ck["end_time"]:
        rc, stt = g.advance(50)
        assert rc == 0, stt
        t = stt["t"]
        phi, c = g.get_state((N, N))[:2]
        assert abs(c.sum() * dx * dx - I0) <= ck["conc_integral_tol"], t
    fs = np.abs(phi).sum() * dx * dx / (1.8 * 1.8)
    print("t", t, "solid fraction", fs, "steps", stt["nst"])
    assert abs(fs - ck["solid_fraction"]) <= ck["solid_fraction_tol"], fs
