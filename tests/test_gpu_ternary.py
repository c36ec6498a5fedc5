"""GPU parity of the three-phase ternary multi-order model against the
oracle (oracle/oracle_ternary.c): closed-form phase concentrations over the
ghost box, the multi-order phase RHS with Moelans' driving force, the
three-phase face diffusivities and the two-species EBS composition RHS
(2-D periodic 40 x 24, 3-D 20 x 12 x 10), then the integrator with several
order parameters and two species step by step (imposed and adaptive steps),
and the reference's Sintering regression replayed on the device.
Tolerances: 1e-14 for the phase concentrations and diffusivities (same
arithmetic), 1e-12 for the right-hand sides, 1e-10 per step for the
integrator fields.
"""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu

GEOMS = {2: (40, 24, 1), 3: (20, 12, 10)}


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _params(ndim, seed):
    p, _, _ = P.sintering_deck()
    p.ndim = ndim
    rng = _rng(seed)
    for a in range(3):
        for k in range(2):
            p.tern_A[a][k] = float(rng.uniform(1e3, 1e4))
            p.tern_ceq[a][k] = float(rng.uniform(0.05, 0.5))
    p.vm_L, p.vm_A, p.vm_B = 1.5e-5, 1.2e-5, 1.8e-5
    p.D_liquid, p.Q0_liquid = 2.0e3, 1.0e4
    p.D_solid, p.Q0_solid = 1.0e4, 2.0e4
    p.D_solid_B, p.Q0_solid_B = 5.0e3, 3.0e4
    return p


@pytest.mark.parametrize("ndim", [2, 3])
def test_ternary_kernels_match_oracle(dev, ndim):
    from oracle import oracle as O
    from ampe_amd import hip as H
    p = _params(ndim, 21 + ndim)
    nx, ny, nz = GEOMS[ndim]
    pt = O.Patch(ndim, (nx, ny, nz), dx=tuple(p.dx))
    box = P.Box()
    box.lo[:] = [0, 0, 0]
    box.hi[:] = [nx - 1, ny - 1, nz - 1 if ndim == 3 else 0]
    lev = H.Level(ndim, [box], 1, periodic=box)
    print(f"ternary kernels ndim {ndim}", flush=True)
    rng = _rng(5 + ndim)
    shape = (pt.nz, pt.ny, pt.nx)
    phi = pt.load(pt.cells(5), rng.uniform(0.02, 1.0, (5,) + shape))
    c = pt.load(pt.cells(2), rng.uniform(0.0, 0.5, (2,) + shape))
    G = pt.G
    ph_o = pt.cells(6)
    O.lib().o_kks_tern3(C.byref(G), C.byref(p), 5, O.ptr(phi), O.ptr(c), O.ptr(ph_o))
    phd, cd = _dev(phi, dev), _dev(c, dev)
    ph = _dev(np.zeros_like(ph_o), dev)
    H.check(H.lib().ampe_kks_ternary3(lev.h, C.byref(p), _p(phd), _p(cd), _p(ph), None))
    assert _rel(ph.cpu().numpy(), ph_o) < 1e-14
    # phase RHS from the same phase concentrations
    r_o = pt.cells(5)
    O.lib().o_phase_rhs_tern3(C.byref(G), C.byref(p), 5, O.ptr(phi), O.ptr(ph_o), None,
                              O.ptr(r_o))
    pho = _dev(ph_o, dev)
    r = _dev(np.zeros_like(r_o), dev)
    H.check(H.lib().ampe_phase_rhs_ternary3(lev.h, C.byref(p), _p(phd), _p(pho), None, _p(r),
                                            None))
    assert _rel(pt.interior(r.cpu().numpy()), pt.interior(r_o)) < 1e-12
    # diffusivities and the composition RHS
    DL, DA, DB = (np.zeros(pt.nf) for _ in range(3))
    O.lib().o_ebs3_coeffs(C.byref(G), C.byref(p), 5, O.ptr(phi), O.ptr(DL), O.ptr(DA), O.ptr(DB))
    dd = [_dev(np.zeros(pt.nf), dev) for _ in range(3)]
    H.check(H.lib().ampe_ebs3_coeffs(lev.h, C.byref(p), _p(phd), *[_p(t) for t in dd], None))
    for t, want in zip(dd, (DL, DA, DB)):
        assert _rel(t.cpu().numpy(), want) < 1e-14
    rc_o = pt.cells(2)
    O.lib().o_conc_rhs_ebs3(C.byref(G), C.byref(p), O.ptr(ph_o), O.ptr(DL), O.ptr(DA), O.ptr(DB),
                            O.ptr(rc_o))
    rc = _dev(np.zeros_like(rc_o), dev)
    dD = [_dev(a, dev) for a in (DL, DA, DB)]
    H.check(H.lib().ampe_conc_rhs_ebs3(lev.h, C.byref(p), _p(pho), *[_p(t) for t in dD], _p(rc),
                                       None))
    assert _rel(pt.interior(rc.cpu().numpy()), pt.interior(rc_o)) < 1e-13


def _small_state(n):
    """A grain of solid A and one of solid B in a liquid, 2-D n x n."""
    from ampe_amd import initial
    ic = initial.make_multi_spheres(n, n, [(n * 0.3, n * 0.5, 0, n * 0.2, "A"),
                                           (n * 0.7, n * 0.5, 0, n * 0.2, "B")],
                                    conc_a=(0.9, 0.05), conc_b=(0.05, 0.9),
                                    conc_out=(0.1, 0.1), width=1.0)
    g = ic["phase"][:, 0].astype(np.float64)
    phi = np.concatenate([g, np.clip(1.0 - g.sum(axis=0), 0.0, 1.0)[None]])
    return phi, ic["conc"][:, 0].astype(np.float64)


@pytest.mark.parametrize("adaptive", [0, 1])
def test_ternary_integrator_matches_oracle(dev, adaptive):
    from oracle import oracle as O
    from ampe_amd import hip as H
    n = 32
    p, _, _ = P.sintering_deck()
    p.norderp, p.norderp_A = 3, 1
    p.tern_ceq[0][0], p.tern_ceq[0][1] = 0.1, 0.1
    p.tern_ceq[1][0], p.tern_ceq[1][1] = 0.9, 0.05
    p.tern_ceq[2][0], p.tern_ceq[2][1] = 0.05, 0.9
    p.D_liquid = 2.0e3
    phi, c = _small_state(n)
    s = P.solver_defaults(dt=2e-5, dt_init=2e-6, adaptive=adaptive, atol=1e-6, rtol=1e-8,
                          maxl=10, max_newton=5)
    so = O.Sim(p, s, n, True)
    so.set_state(phi, c, None)
    box = H.boxes_of((n, n), 2)
    sg = H.Sim(p, s, [box], periodic=box)
    sg.set_state(phi, c, None)
    for step in range(6):
        print(f"ternary integrator adaptive {adaptive} step {step}", flush=True)
        ro, sto = so.advance(1)
        rg, stg = sg.advance(1)
        assert ro == rg == 0, (step, sto, stg)
        for k in ("nst", "nni", "nli", "ncfn", "nlcf", "netf"):
            assert sto[k] == stg[k], (step, k, sto, stg)
        # adaptive steps come from error norms: equal to the fields' rounding
        assert abs(sto["t"] - stg["t"]) <= (1e-9 if adaptive else 1e-12) * sto["t"]
        po, co, _, clo, cso = so.get_state()
        pg, cg, _, clg, csg = sg.get_state((n, n))
        assert _rel(pg, po) < 1e-10, (step, _rel(pg, po))
        assert _rel(cg, co) < 1e-10, (step, _rel(cg, co))
        assert _rel(clg, clo) < 1e-10 and _rel(csg, cso) < 1e-10


def test_sintering_regression_device(dev):
    """The reference's Sintering regression (tests/Sintering/test2d.py) on the
    device at Newton atol 1e-6: the largest and smallest order-parameter
    volume fractions at t = 0.006 s are 0.586 and 0.118 (+- 0.001); the
    oracle's run of the same deck agrees to 1e-6."""
    from oracle import oracle as O
    from ampe_amd import hip as H
    p, (phi, c), chk = P.sintering_deck()
    s = P.solver_defaults(dt=5e-4, dt_init=1e-6, adaptive=1, atol=1e-6, rtol=1e-8, maxl=10,
                          max_newton=3)
    box = H.boxes_of((64, 64), 2)
    sg = H.Sim(p, s, [box], periodic=box)
    sg.set_state(phi, c, None)
    so = O.Sim(p, s, 64, True)
    so.set_state(phi, c, None)
    out = []
    for name, sim, get in (("device", sg, lambda: sg.get_state((64, 64))[0]),
                           ("oracle", so, lambda: so.get_state()[0])):
        print(f"sintering {name}", flush=True)
        t, n = 0.0, 0
        while t <= chk["end_time"] and n < 2000:
            rc, st = sim.advance(1)
            assert rc == 0, st
            t, n = st["t"], n + 1
        ph = get()
        out.append((ph.reshape(5, -1).mean(axis=1), t, n))
    vol = out[0][0]
    print(f"device volumes {vol} at t {out[0][1]} after {out[0][2]} steps", flush=True)
    assert abs(vol.max() - chk["vmax"]) <= chk["tol"], vol
    assert abs(vol.min() - chk["vmin"]) <= chk["tol"], vol
    assert out[0][2] == out[1][2]
    assert np.max(np.abs(out[0][0] - out[1][0])) < 1e-6, (out[0][0], out[1][0])
