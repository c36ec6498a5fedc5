"""GPU parity: every HIP kernel against the FP64 C oracle on identical seeded
inputs, through the C ABI.  Tolerances: 1e-13 relative for kernels with no
transcendental beyond sqrt (same operation order, no FMA contraction), 1e-12
where log/exp/atan/acos/cos/sin/tanh enter (device libm vs glibc differ by an
ulp), bit-exact for integer tags.
"""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu

GEOMS = {2: (40, 24, 1), 3: (20, 12, 10)}


def _setup(ndim, p=None):
    """Oracle patch (with the model's dx) and the matching device level."""
    from oracle import oracle as O
    from ampe_amd import hip as H
    nx, ny, nz = GEOMS[ndim]
    dx = tuple(p.dx) if p is not None else (1.0, 1.0, 1.0)
    pt = O.Patch(ndim, (nx, ny, nz), dx=dx)
    box = P.Box()
    box.lo[:] = [0, 0, 0]
    box.hi[:] = [nx - 1, ny - 1, nz - 1 if ndim == 3 else 0]
    lev = H.Level(ndim, [box], 1, periodic=box)
    assert lev.ncg == pt.ncg and lev.nside_tot == pt.nf
    return O, H, pt, lev


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _fields(pt, seed=7):
    rng = _rng(seed)
    shape = (pt.nz, pt.ny, pt.nx)
    phi = pt.load(pt.cells(), rng.uniform(0.0, 1.0, shape))
    qv = rng.normal(size=(4,) + shape)
    qv /= np.sqrt((qv ** 2).sum(0))
    q = pt.load(pt.cells(4), qv)
    c = pt.load(pt.cells(), rng.uniform(0.3, 0.5, shape))
    cl = pt.load(pt.cells(), rng.uniform(0.2, 0.6, shape))
    cs = pt.load(pt.cells(), rng.uniform(0.2, 0.6, shape))
    return phi, q, c, cl, cs


def _dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _params(ndim, kind):
    if kind == "calphad":
        return P.polycrystal_params(ndim=ndim)
    if kind == "dilute":
        return P.dilute_params(ndim=ndim)
    p = P.two_grain_params(ndim=ndim)
    if kind == "iso":
        p.flux_type = P.FLUX_ISOTROPIC
    elif kind == "aniso":
        p.flux_type = P.FLUX_ANISOTROPIC
        p.aniso_nu = 0.05
        p.aniso_knumber = 4
    return p


@pytest.mark.parametrize("ndim", [2, 3])
def test_quat_grad(dev, ndim):
    p = _params(ndim, "simple")
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt)
    gm_o = pt.cells()
    n2_o = np.zeros(pt.nf)
    O.lib().o_quat_grad_modulus(C.byref(pt.G), C.byref(p), O.ptr(q), O.ptr(gm_o))
    O.lib().o_quat_face_norm2(C.byref(pt.G), C.byref(p), O.ptr(q), O.ptr(n2_o))
    qd = _dev(q, dev)
    gm = _dev(np.zeros_like(gm_o), dev)
    n2 = _dev(np.zeros(pt.nf), dev)
    H.check(H.lib().ampe_quat_grad(lev.h, C.byref(p), _p(qd), _p(gm), _p(n2), None))
    assert _rel(pt.interior(gm.cpu().numpy()), pt.interior(gm_o)) < 1e-13
    assert _rel(n2.cpu().numpy(), n2_o) < 1e-13


@pytest.mark.parametrize("ndim,kind", [(2, "simple"), (2, "iso"), (2, "aniso"), (3, "simple"),
                                       (3, "aniso"), (2, "calphad"), (3, "calphad"),
                                       (2, "dilute"), (3, "dilute")])
def test_phase_rhs(dev, ndim, kind):
    p = _params(ndim, kind)
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=11)
    r_o = pt.cells()
    O.lib().o_phase_rhs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(q), O.ptr(cl),
                        O.ptr(cs), O.ptr(r_o))
    ds = [_dev(a, dev) for a in (phi, q, cl, cs)]
    r = _dev(np.zeros_like(r_o), dev)
    H.check(H.lib().ampe_phase_rhs(lev.h, C.byref(p), *[_p(t) for t in ds], _p(r), None))
    tol = 1e-12 if kind in ("aniso", "calphad", "dilute") else 1e-13
    assert _rel(pt.interior(r.cpu().numpy()), pt.interior(r_o)) < tol


@pytest.mark.parametrize("ndim", [2, 3])
def test_quat_rhs_and_project(dev, ndim):
    p = _params(ndim, "simple")
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=13)
    n2 = np.zeros(pt.nf)
    O.lib().o_quat_face_norm2(C.byref(pt.G), C.byref(p), O.ptr(q), O.ptr(n2))
    n2 *= _rng(3).uniform(0.5, 1.5, n2.shape)  # lagged norms differ from current
    r_o = pt.cells(4)
    O.lib().o_quat_rhs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(q), O.ptr(n2), O.ptr(r_o))
    phd, qd, n2d = _dev(phi, dev), _dev(q, dev), _dev(n2, dev)
    r = _dev(np.zeros_like(r_o), dev)
    H.check(H.lib().ampe_quat_rhs(lev.h, C.byref(p), _p(phd), _p(qd), _p(n2d), _p(r), None))
    assert _rel(pt.interior(r.cpu().numpy()), pt.interior(r_o)) < 1e-12
    # two passes (face coefficients, then the RHS from them): the same values
    fcd = _dev(np.zeros(pt.nf), dev)
    r2 = _dev(np.zeros_like(r_o), dev)
    H.check(H.lib().ampe_quat_face_coefs(lev.h, C.byref(p), _p(phd), _p(n2d), _p(fcd), None))
    H.check(H.lib().ampe_quat_rhs_fc(lev.h, C.byref(p), _p(phd), _p(qd), _p(fcd), _p(r2), None))
    assert np.array_equal(pt.interior(r2.cpu().numpy()), pt.interior(r.cpu().numpy()))
    # projection
    err = pt.cells(4) + _rng(5).normal(size=pt.cells(4).shape)
    corr_o, err_o = pt.cells(4), err.copy()
    O.lib().o_quat_project(C.byref(pt.G), C.byref(p), O.ptr(q), O.ptr(corr_o), O.ptr(err_o))
    corr, errd = _dev(pt.cells(4), dev), _dev(err, dev)
    H.check(H.lib().ampe_quat_project(lev.h, C.byref(p), _p(qd), _p(corr), _p(errd), None))
    assert _rel(pt.interior(corr.cpu().numpy()), pt.interior(corr_o)) < 1e-13
    assert _rel(pt.interior(errd.cpu().numpy()), pt.interior(err_o)) < 1e-13


@pytest.mark.parametrize("ndim", [2, 3])
def test_conc(dev, ndim):
    p = _params(ndim, "simple")
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=17)
    Dc_o, Dp_o = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_conc_coeffs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(cl), O.ptr(cs),
                          O.ptr(Dc_o), O.ptr(Dp_o))
    r_o = pt.cells()
    O.lib().o_conc_rhs(C.byref(pt.G), C.byref(p), O.ptr(c), O.ptr(phi), O.ptr(Dc_o),
                       O.ptr(Dp_o), O.ptr(r_o))
    phd, cld, csd, cd = (_dev(a, dev) for a in (phi, cl, cs, c))
    Dc, Dp = _dev(np.zeros(pt.nf), dev), _dev(np.zeros(pt.nf), dev)
    H.check(H.lib().ampe_conc_coeffs(lev.h, C.byref(p), _p(phd), _p(cld), _p(csd), _p(Dc),
                                     _p(Dp), None))
    assert _rel(Dc.cpu().numpy(), Dc_o) < 1e-12
    assert _rel(Dp.cpu().numpy(), Dp_o) < 1e-12
    r = _dev(np.zeros_like(r_o), dev)
    Dco, Dpo = _dev(Dc_o, dev), _dev(Dp_o, dev)
    H.check(H.lib().ampe_conc_rhs(lev.h, C.byref(p), _p(cd), _p(phd), _p(Dco), _p(Dpo), _p(r),
                                  None))
    assert _rel(pt.interior(r.cpu().numpy()), pt.interior(r_o)) < 1e-13


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("interp", ["l", "p"])
def test_conc_ebs(dev, ndim, interp):
    """EBS composition form: phase-weighted face diffusivities and the
    divergence of D_L grad cL + D_A grad cS."""
    p = _params(ndim, "simple")
    p.diff_interp = interp.encode()
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=29)
    DL_o, DA_o = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_ebs_coeffs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(DL_o), O.ptr(DA_o))
    r_o = pt.cells()
    O.lib().o_conc_rhs_ebs(C.byref(pt.G), C.byref(p), O.ptr(cl), O.ptr(cs), O.ptr(DL_o),
                           O.ptr(DA_o), O.ptr(r_o))
    phd, cld, csd = (_dev(a, dev) for a in (phi, cl, cs))
    DL, DA = _dev(np.zeros(pt.nf), dev), _dev(np.zeros(pt.nf), dev)
    H.check(H.lib().ampe_conc_coeffs_ebs(lev.h, C.byref(p), _p(phd), _p(DL), _p(DA), None))
    assert _rel(DL.cpu().numpy(), DL_o) < 1e-13
    assert _rel(DA.cpu().numpy(), DA_o) < 1e-13
    r = _dev(np.zeros_like(r_o), dev)
    H.check(H.lib().ampe_conc_rhs_ebs(lev.h, C.byref(p), _p(cld), _p(csd), _p(DL), _p(DA),
                                      _p(r), None))
    assert _rel(pt.interior(r.cpu().numpy()), pt.interior(r_o)) < 1e-13


@pytest.mark.parametrize("ndim", [2, 3])
def test_conc_ebs_isotropic(dev, ndim):
    """EBS composition form with the isotropic stencil (add_flux_iso): the
    device kernel, which evaluates every face diffusivity it touches from phi,
    against the oracle's restatement."""
    p = _params(ndim, "simple")
    p.diff_interp = b"p"
    p.ebs_stencil = P.EBS_ISOTROPIC
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=31)
    r_o = pt.cells()
    O.lib().o_conc_rhs_ebs_iso(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(cl), O.ptr(cs),
                               O.ptr(r_o))
    phd, cld, csd = (_dev(a, dev) for a in (phi, cl, cs))
    r = _dev(np.zeros_like(r_o), dev)
    H.check(H.lib().ampe_conc_rhs_ebs_iso(lev.h, C.byref(p), _p(phd), _p(cld), _p(csd), _p(r),
                                          None))
    assert _rel(pt.interior(r.cpu().numpy()), pt.interior(r_o)) < 1e-13


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("layout", ["periodic", "bounded", "bounded_two_patches"])
def test_conc_ebs_order4(dev, ndim, layout):
    """EBS composition form with the fourth-order stencil (add_flux_4th) on
    levels with two ghost cells: periodic, and bounded (zero-slope ghosts,
    second-order faces on the physical boundary) as one patch or as two
    patches split along x, whose shared faces must stay fourth order."""
    from oracle import oracle as O
    from ampe_amd import hip as H
    p = _params(ndim, "simple")
    p.diff_interp = b"p"
    p.conc_rhs_form = P.CONC_RHS_EBS
    p.ebs_stencil = P.EBS_ORDER4
    nx, ny, nz = GEOMS[ndim]
    pt = O.Patch(ndim, (nx, ny, nz), g=2, dx=tuple(p.dx))
    pt.G.periodic = 1 if layout == "periodic" else 0
    rng = _rng(37)
    shape = (pt.nz, pt.ny, pt.nx)
    vals = [rng.uniform(0.0, 1.0, shape), rng.uniform(0.2, 0.6, shape),
            rng.uniform(0.2, 0.6, shape)]
    phi, cl, cs = (pt.cells() for _ in range(3))
    for a, v in zip((phi, cl, cs), vals):
        pt.interior(a)[...] = v[None] if ndim == 3 else v[:1][None]
        O.lib().o_ghost_fill(C.byref(pt.G), O.ptr(a), 1)
    DL_o, DA_o = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_ebs_coeffs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(DL_o), O.ptr(DA_o))
    r_o = pt.cells()
    O.lib().o_conc_rhs_ebs_4th(C.byref(pt.G), C.byref(p), O.ptr(cl), O.ptr(cs), O.ptr(DL_o),
                               O.ptr(DA_o), O.ptr(r_o))
    hi = [nx - 1, ny - 1, nz - 1 if ndim == 3 else 0]
    if layout == "bounded_two_patches":
        b0, b1 = P.Box(), P.Box()
        b0.lo[:] = [0, 0, 0]
        b0.hi[:] = [nx // 2 - 1] + hi[1:]
        b1.lo[:] = [nx // 2, 0, 0]
        b1.hi[:] = hi
        boxes = [b0, b1]
    else:
        b = P.Box()
        b.lo[:] = [0, 0, 0]
        b.hi[:] = hi
        boxes = [b]
    lev = H.Level(ndim, boxes, 2, periodic=boxes[0] if layout == "periodic" else None)
    dom = [v if ndim == 3 else v[0] for v in vals]
    phd, cld, csd = (_dev(lev.to_ghosted(v, ghost=2), dev) for v in dom)
    for t in (phd, cld, csd):
        H.check(H.lib().ampe_ghost_fill(lev.h, _p(t), 1, None))
    DL = _dev(np.zeros(lev.nside_tot), dev)
    DA = _dev(np.zeros(lev.nside_tot), dev)
    H.check(H.lib().ampe_conc_coeffs_ebs(lev.h, C.byref(p), _p(phd), _p(DL), _p(DA), None))
    r = _dev(np.zeros(lev.ncg), dev)
    H.check(H.lib().ampe_conc_rhs_ebs_4th(lev.h, C.byref(p), _p(cld), _p(csd), _p(DL), _p(DA),
                                          _p(r), None))
    rh = r.cpu().numpy()
    got = np.zeros(shape)
    for q, bx in enumerate(boxes):
        blk = lev.ghost_view(rh, q)
        blk = blk if ndim == 3 else blk[None]
        n = [bx.hi[d] - bx.lo[d] + 1 for d in range(3)]
        got[bx.lo[2]:bx.lo[2] + n[2], :, bx.lo[0]:bx.lo[0] + n[0]] = \
            blk[2 * (ndim == 3):2 * (ndim == 3) + n[2], 2:2 + n[1], 2:2 + n[0]]
    assert _rel(got, pt.interior(r_o)[0]) < 1e-13
    # a level with one ghost cell is refused
    lev1 = H.Level(ndim, boxes, 1, periodic=boxes[0] if layout == "periodic" else None)
    assert H.lib().ampe_conc_rhs_ebs_4th(lev1.h, C.byref(p), _p(cld), _p(csd), _p(DL), _p(DA),
                                         _p(r), None) != 0


@pytest.mark.parametrize("ndim,kind", [(2, "simple"), (3, "calphad"), (3, "dilute")])
def test_kks(dev, ndim, kind):
    p = _params(ndim, kind)
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=19)
    clr, csr = c.copy(), c.copy()
    cl_o, cs_o = pt.cells(), pt.cells()
    it_o = O.lib().o_kks_solve(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(c), O.ptr(clr),
                               O.ptr(csr), O.ptr(cl_o), O.ptr(cs_o))
    assert it_o >= 0
    ds = [_dev(a, dev) for a in (phi, c, clr, csr)]
    cld, csd = _dev(pt.cells(), dev), _dev(pt.cells(), dev)
    it = C.c_int(0)
    rc = H.check(H.lib().ampe_kks_solve(lev.h, C.byref(p), *[_p(t) for t in ds], _p(cld),
                                        _p(csd), C.byref(it), None))
    assert rc == 0 and it.value == it_o
    assert _rel(cld.cpu().numpy(), cl_o) < 1e-12
    assert _rel(csd.cpu().numpy(), cs_o) < 1e-12


@pytest.mark.parametrize("ndim", [2, 3])
def test_tags_bit_exact(dev, ndim):
    import torch
    p = _params(ndim, "simple")
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=23)
    tags_o = np.zeros(pt.N, dtype=np.int32)
    O.lib().o_tag_cells(C.byref(pt.G), C.byref(p), O.ptr(phi), 0.3, O.iptr(tags_o))
    phd = _dev(phi, dev)
    tags = torch.zeros(pt.N, dtype=torch.int32, device=dev)
    H.check(H.lib().ampe_tag_cells(lev.h, C.byref(p), _p(phd), 0.3, _p(tags), None))
    assert np.array_equal(tags.cpu().numpy(), tags_o)
    assert 0 < tags_o.sum() < pt.N


@pytest.mark.parametrize("variant", ["inf", "finite", "antitrap"])
def test_kim_phase_mobility(dev, variant):
    """Kim's KKS phase mobility (three reference variants): the mobility field,
    the phase preconditioner coefficient C built from it and the phase RHS that
    multiplies by it, GPU vs oracle (dilute CuNi model, 3-D)."""
    p = P.dilute_params(ndim=3, phi_mobility_type=P.PHI_MOB_KIM, D_liquid=1.0e3,
                        Q0_liquid=0.0, energy_interp=b"h")
    if variant != "inf":
        p.kim_interface_mobility = P.kim_interface_mobility(p.vm_L, p.dilute_Tm, p.dilute_m,
                                                            p.dilute_keq, 2.0e-5)
        p.kim_antitrapping = 1 if variant == "antitrap" else 0
    O, H, pt, lev = _setup(3, p)
    phi, q, c, cl, cs = _fields(pt, seed=41)
    M_o = pt.cells()
    O.lib().o_phase_mobility_field(C.byref(pt.G), C.byref(p), O.ptr(cl), O.ptr(cs), O.ptr(M_o))
    cld, csd, phd = _dev(cl, dev), _dev(cs, dev), _dev(phi, dev)
    Md = _dev(pt.cells(), dev)
    H.check(H.lib().ampe_phase_mobility(lev.h, C.byref(p), _p(cld), _p(csd), _p(Md), None))
    Mg = pt.interior(Md.cpu().numpy())
    assert _rel(Mg, pt.interior(M_o)) < 1e-14
    assert np.all(np.isfinite(Mg)) and np.all(Mg > 0)
    if variant == "finite":
        assert np.ptp(Mg) == 0.0  # 1 / alpha, independent of the concentrations
    gamma = 2.0e-6
    C_o = pt.cells()
    O.lib().o_phase_fac_C(C.byref(pt.G), C.byref(p), gamma, O.ptr(phi), O.ptr(M_o), O.ptr(C_o))
    Cd = _dev(pt.cells(), dev)
    H.check(H.lib().ampe_fac_phase_coeffs(lev.h, C.byref(p), gamma, _p(phd), _p(Md), _p(Cd),
                                          None))
    assert _rel(pt.interior(Cd.cpu().numpy()), pt.interior(C_o)) < 1e-14
    r_o = pt.cells()
    O.lib().o_phase_rhs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(q), O.ptr(cl), O.ptr(cs),
                        O.ptr(r_o))
    ds = [_dev(a, dev) for a in (phi, q, cl, cs)]
    r = _dev(np.zeros_like(r_o), dev)
    H.check(H.lib().ampe_phase_rhs(lev.h, C.byref(p), *[_p(t) for t in ds], _p(r), None))
    assert _rel(pt.interior(r.cpu().numpy()), pt.interior(r_o)) < 1e-12


@pytest.mark.parametrize("ndim", [2, 3])
def test_fac_smooth_residual(dev, ndim):
    p = _params(ndim, "simple")
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=29)
    gamma = 1.0e-3
    C_o = pt.cells()
    O.lib().o_phase_fac_C(C.byref(pt.G), C.byref(p), gamma, O.ptr(phi), None, O.ptr(C_o))
    phd = _dev(phi, dev)
    Cd = _dev(pt.cells(), dev)
    H.check(H.lib().ampe_fac_phase_coeffs(lev.h, C.byref(p), gamma, _p(phd), None, _p(Cd),
                                          None))
    assert _rel(pt.interior(Cd.cpu().numpy()), pt.interior(C_o)) < 1e-14
    rhs = pt.load(pt.cells(), _rng(31).normal(size=(pt.nz, pt.ny, pt.nx)))
    Dside = -gamma * _rng(37).uniform(1.0, 5.0, pt.nf)
    for variant in ("phase", "conc"):
        op_o = O.EllipticOp()
        op_h = H.EllipticOp()
        dx = (C.c_double * 3)(1.0, 1.0, 1.0)
        Dd = _dev(Dside, dev)
        if variant == "phase":
            vals = dict(D=None, D_const=-gamma * p.epsilon_phi ** 2, C_const=1.0,
                        M_const=p.phi_mobility)
            op_o.C, op_h.C = C_o.ctypes.data, Cd.data_ptr()
        else:
            vals = dict(D_const=0.0, C_const=1.0, M_const=1.0)
            op_o.D, op_h.D = Dside.ctypes.data, Dd.data_ptr()
            op_o.C = op_h.C = None
        for k, v in vals.items():
            if v is not None:
                setattr(op_o, k, v)
                setattr(op_h, k, v)
        op_o.dx = op_h.dx = dx
        u_o = pt.cells()
        flux = np.zeros(pt.nf)
        m_o = O.lib().o_fac_smooth(C.byref(pt.G), C.byref(op_o), O.ptr(rhs), O.ptr(u_o), 3,
                                   -1.0, O.ptr(flux))
        rhsd, ud = _dev(rhs, dev), _dev(pt.cells(), dev)
        mres = C.c_double(0)
        H.check(H.lib().ampe_fac_smooth(lev.h, C.byref(op_h), _p(rhsd), _p(ud), 3, -1.0,
                                        C.byref(mres), None))
        assert _rel(pt.interior(ud.cpu().numpy()), pt.interior(u_o)) < 1e-13
        assert abs(mres.value - m_o) <= 1e-13 * abs(m_o)
        res_o = pt.cells()
        rms_o = O.lib().o_fac_residual(C.byref(pt.G), C.byref(op_o), O.ptr(rhs), O.ptr(u_o),
                                       O.ptr(res_o), O.ptr(flux))
        resd = _dev(pt.cells(), dev)
        rms = C.c_double(0)
        H.check(H.lib().ampe_fac_residual(lev.h, C.byref(op_h), _p(rhsd), _p(ud), _p(resd),
                                          C.byref(rms), None))
        assert _rel(pt.interior(resd.cpu().numpy()), pt.interior(res_o)) < 1e-12
        assert abs(rms.value - rms_o) <= 1e-12 * rms_o


@pytest.mark.parametrize("ndim", [2, 3])
def test_reductions(dev, ndim):
    O, H, pt, lev = _setup(ndim)
    rng = _rng(41)
    x = rng.normal(size=pt.cells(6).shape)
    y = rng.normal(size=pt.cells(6).shape)
    w = rng.uniform(0.5, 2.0, size=pt.cells(6).shape)
    d_o = O.lib().o_dot(C.byref(pt.G), 6, O.ptr(x), O.ptr(y))
    w_o = O.lib().o_wrms(C.byref(pt.G), 6, O.ptr(x), O.ptr(w))
    xd, yd, wd = _dev(x, dev), _dev(y, dev), _dev(w, dev)
    out = C.c_double(0)
    H.check(H.lib().ampe_vec_dot(lev.h, 6, _p(xd), _p(yd), C.byref(out), None))
    assert abs(out.value - d_o) <= 1e-12 * np.abs(x * y).sum()
    H.check(H.lib().ampe_vec_wrms(lev.h, 6, _p(xd), _p(wd), C.byref(out), None))
    assert abs(out.value - w_o) <= 1e-13 * w_o


@pytest.mark.parametrize("ndim", [2, 3])
def test_multipatch_level_matches_single(dev, ndim):
    """Uneven patches of one level in one launch give the single-patch result
    (ghost gather between patches and across the periodic boundary)."""
    import torch
    p = _params(ndim, "aniso" if ndim == 2 else "simple")
    O, H, pt, lev = _setup(ndim, p)
    phi, q, c, cl, cs = _fields(pt, seed=43)
    nx, ny, nz = GEOMS[ndim]
    cuts = [0, 13, 29, nx] if ndim == 2 else [0, 7, nx]
    boxes = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        for (j0, j1) in ([(0, 9), (10, ny - 1)] if ndim == 2 else [(0, ny - 1)]):
            bx = P.Box()
            bx.lo[:] = [a, j0, 0]
            bx.hi[:] = [b - 1, j1, nz - 1 if ndim == 3 else 0]
            boxes.append(bx)
    dom = P.Box()
    dom.lo[:] = [0, 0, 0]
    dom.hi[:] = [nx - 1, ny - 1, nz - 1 if ndim == 3 else 0]
    ml = H.Level(ndim, boxes, 1, periodic=dom)

    def scatter(arr_g, ncomp):
        """single-patch ghosted array -> multi-patch layout (interior only)"""
        out = np.zeros((ncomp, ml.ncg))
        src = arr_g.reshape(ncomp, -1)
        sy, sz = nx + 2, (nx + 2) * (ny + 2)
        gz = 1 if ndim == 3 else 0
        for b, off in zip(boxes, ml.cell_off):
            bnx, bny = b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1
            bnz = (b.hi[2] - b.lo[2] + 1) if ndim == 3 else 1
            bsy, bsz = bnx + 2, (bnx + 2) * (bny + 2)
            for k in range(bnz):
                for j in range(bny):
                    gi = (b.lo[0] + 1) + (b.lo[1] + j + 1) * sy + (b.lo[2] + k + gz) * sz
                    li = off + 1 + (j + 1) * bsy + (k + gz) * bsz
                    out[:, li:li + bnx] = src[:, gi:gi + bnx]
        return out

    def gather(arr_m, ncomp):
        out = np.zeros((ncomp, pt.ncg))
        sy, sz = nx + 2, (nx + 2) * (ny + 2)
        gz = 1 if ndim == 3 else 0
        for b, off in zip(boxes, ml.cell_off):
            bnx, bny = b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1
            bnz = (b.hi[2] - b.lo[2] + 1) if ndim == 3 else 1
            bsy, bsz = bnx + 2, (bnx + 2) * (bny + 2)
            for k in range(bnz):
                for j in range(bny):
                    gi = (b.lo[0] + 1) + (b.lo[1] + j + 1) * sy + (b.lo[2] + k + gz) * sz
                    li = off + 1 + (j + 1) * bsy + (k + gz) * bsz
                    out[:, gi:gi + bnx] = arr_m.reshape(ncomp, -1)[:, li:li + bnx]
        return out

    # single patch reference on the GPU
    ds = [_dev(a, dev) for a in (phi, q, cl, cs)]
    r1 = _dev(pt.cells(), dev)
    H.check(H.lib().ampe_phase_rhs(lev.h, C.byref(p), *[_p(t) for t in ds], _p(r1), None))
    # multi patch: interior only, ghosts come from the gather
    dm = [_dev(scatter(a, n), dev) for a, n in ((phi, 1), (q, 4), (cl, 1), (cs, 1))]
    H.check(H.lib().ampe_ghost_fill(ml.h, _p(dm[0]), 1, None))
    H.check(H.lib().ampe_ghost_fill(ml.h, _p(dm[1]), 4, None))
    rm = torch.zeros(ml.ncg, dtype=torch.float64, device=dev)
    H.check(H.lib().ampe_phase_rhs(ml.h, C.byref(p), *[_p(t) for t in dm], _p(rm), None))
    got = gather(rm.cpu().numpy(), 1).reshape(pt.gshape)
    ref = r1.cpu().numpy().reshape(pt.gshape)
    assert np.array_equal(pt.interior(got), pt.interior(ref))


def _two_grain(n, ndim=2):
    p = P.two_grain_params(ndim=ndim)
    p.quad_ceq[0], p.quad_ceq[1] = 0.40, 0.45
    phi, c, q = P.two_grain_state(n=n, radius=n / 4, ndim=ndim)
    return p, phi, c, q


@pytest.mark.parametrize("gstype", [P.GS_MODIFIED, P.GS_CLASSICAL])
def test_integrator_fp64_matches_oracle(dev, gstype):
    from oracle import oracle as O
    from ampe_amd import hip as H
    n = 48
    p, phi, c, q = _two_grain(n)
    s = P.solver_defaults(maxl=10, max_newton=5, gstype=gstype)
    osim = O.Sim(p, s, n)
    osim.set_state(phi, c, q)
    box = H.boxes_of((n, n), 2)
    gsim = H.Sim(p, s, [box], periodic=box)
    gsim.set_state(phi, c, q)
    for step in range(6):
        ro, sto = osim.advance(1)
        rg, stg = gsim.advance(1)
        assert ro == rg == 0
        fo = osim.get_state()
        fg = gsim.get_state((n, n))
        for name, a, b in zip(("phi", "c", "q"), fg[:3], fo[:3]):
            e = _rel(a, b)
            assert e < 1e-10, (step, name, e, sto, stg)
        assert sto["nni"] == stg["nni"] and sto["nli"] == stg["nli"], (sto, stg)


def test_config1_two_grain_20_steps_matches_oracle(dev):
    """The plain-method case run in full: 2-D 128^2 periodic, two grains
    (radius 32, tanh width 4), quadratic KKS with c = 0.40 / 0.45 plus noise,
    isotropic phase flux, 20 BDF steps at dt = 1e-3, FP64 red-black block
    solves with 2 sweeps.  The step is cut on convergence failures here, and the
    GPU must take the same cuts: per step, every field within 1e-10 relative L2
    of the oracle, identical Newton / Krylov / failure counters and time."""
    from oracle import oracle as O
    from ampe_amd import hip as H
    n = 128
    p = P.two_grain_params(ndim=2, flux_type=P.FLUX_ISOTROPIC)
    p.quad_ceq[0], p.quad_ceq[1] = 0.40, 0.45
    phi, c, q = P.two_grain_state(n=n)
    s = P.solver_defaults(dt=1e-3, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8,
                          fac_sweeps=2, quat_sweeps=2)
    osim = O.Sim(p, s, n)
    osim.set_state(phi, c, q)
    box = H.boxes_of((n, n), 2)
    gsim = H.Sim(p, s, [box], periodic=box)
    gsim.set_state(phi, c, q)
    for step in range(20):
        ro, sto = osim.advance(1)
        rg, stg = gsim.advance(1)
        assert ro == rg == 0, (step, sto, stg)
        fo = osim.get_state()
        fg = gsim.get_state((n, n))
        for name, a, b in zip(("phi", "c", "q"), fg[:3], fo[:3]):
            e = _rel(a, b)
            assert e < 1e-10, (step, name, e, sto, stg)
        for k in ("nst", "nni", "nli", "ncfn", "nlcf"):
            assert sto[k] == stg[k], (step, k, sto, stg)
        assert sto["t"] == stg["t"], (step, sto["t"], stg["t"])


@pytest.mark.parametrize("model,gstype", [("calphad", P.GS_MODIFIED),
                                          ("calphad", P.GS_CLASSICAL),
                                          ("dilute_kim", P.GS_MODIFIED),
                                          ("calphad_ebs", P.GS_MODIFIED),
                                          ("calphad_ebs_iso", P.GS_MODIFIED),
                                          ("calphad_ebs_4th", P.GS_MODIFIED)])
def test_integrator_3d_kks_matches_oracle(dev, model, gstype):
    """3-D polycrystal with a KKS composition model, FP64 GPU integrator against
    the oracle step by step: binary CALPHAD (Newton inside every RHS
    evaluation) with both Gram-Schmidt variants and with the EBS composition
    flux, and the dilute solution with Kim's phase mobility (mobility field in
    the RHS and the preconditioner)."""
    from oracle import oracle as O
    from ampe_amd import hip as H
    n = 16
    if model.startswith("calphad"):
        p = P.polycrystal_params(seed=1002)
        if model.startswith("calphad_ebs"):
            p.conc_rhs_form = P.CONC_RHS_EBS
        if model == "calphad_ebs_iso":
            p.ebs_stencil = P.EBS_ISOTROPIC
        if model == "calphad_ebs_4th":
            p.ebs_stencil = P.EBS_ORDER4
        dt = 1e-7
    else:
        p = P.dilute_params(ndim=3, phi_mobility_type=P.PHI_MOB_KIM, D_liquid=1.0e3,
                            Q0_liquid=0.0)
        dt = 1e-9
    phi, c, q = P.polycrystal_state(n=n, seed=1002, ngrains=4)
    s = P.solver_defaults(dt=dt, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8, gstype=gstype)
    osim = O.Sim(p, s, n)
    osim.set_state(phi, c, q)
    box = H.boxes_of((n, n, n), 3)
    gsim = H.Sim(p, s, [box], periodic=box)
    gsim.set_state(phi, c, q)
    for step in range(3):
        ro, sto = osim.advance(1)
        rg, stg = gsim.advance(1)
        assert ro == rg == 0
        fo = osim.get_state()
        fg = gsim.get_state((n, n, n))
        for name, a, b in zip(("phi", "c", "q", "cl", "cs"), fg, fo):
            e = _rel(a, b)
            assert e < 1e-10, (step, name, e, sto, stg)
        assert sto["nni"] == stg["nni"] and sto["nli"] == stg["nli"], (sto, stg)


def test_integrator_mixed_precision_close_to_fp64(dev):
    from ampe_amd import hip as H
    n = 48
    p, phi, c, q = _two_grain(n)
    res = {}
    for prec in (P.FP64, P.FAC_FP32):
        s = P.solver_defaults(maxl=10, max_newton=5, atol=1e-8, rtol=1e-10, precision=prec)
        box = H.boxes_of((n, n), 2)
        sim = H.Sim(p, s, [box], periodic=box)
        sim.set_state(phi, c, q)
        rc, st = sim.advance(5)
        assert rc == 0
        res[prec] = (sim.get_state((n, n)), st)
    (a, sa), (b, sb) = res[P.FP64], res[P.FAC_FP32]
    assert abs(sa["t"] - sb["t"]) < 1e-15
    for name, x, y in zip(("phi", "c", "q"), a[:3], b[:3]):
        assert _rel(y, x) < 1e-6, (name, _rel(y, x))


@pytest.mark.parametrize("case", ["two_grain_2d", "calphad_3d"])
def test_adaptive_integrator_matches_oracle(dev, case):
    """Adaptive BDF (local error test, step and order selection, CVODE's
    fixed-leading-coefficient form) on the device against the oracle step by
    step: identical step sizes, orders, error-test and convergence failures
    and Newton / Krylov counters, every field within 1e-10."""
    from oracle import oracle as O
    from ampe_amd import hip as H
    if case == "two_grain_2d":
        n = 48
        p, phi, c, q = _two_grain(n)
        shape = (n, n)
        s = P.solver_defaults(dt=1e-2, dt_init=1e-4, maxl=10, max_newton=4, atol=1e-5,
                              rtol=1e-7, adaptive=1)
        nsteps = 12
    else:
        n = 16
        p = P.polycrystal_params(seed=1002)
        phi, c, q = P.polycrystal_state(n=n, seed=1002, ngrains=4)
        shape = (n, n, n)
        s = P.solver_defaults(dt=1e-6, dt_init=1e-9, maxl=10, max_newton=4, atol=1e-6,
                              rtol=1e-8, adaptive=1)
        nsteps = 8
    osim = O.Sim(p, s, shape[::-1] if len(shape) == 3 else (n, n, 1))
    osim.set_state(phi, c, q)
    box = H.boxes_of(shape, len(shape))
    gsim = H.Sim(p, s, [box], periodic=box)
    gsim.set_state(phi, c, q)
    qs = set()
    for step in range(nsteps):
        ro, sto = osim.advance(1)
        rg, stg = gsim.advance(1)
        assert ro == rg == 0, (step, sto, stg)
        for k in ("nst", "nni", "nli", "ncfn", "netf", "q"):
            assert sto[k] == stg[k], (step, k, sto, stg)
        # the step size follows the local error estimate, a difference of
        # nearly equal states, so it carries the fields' rounding (~1e-11)
        assert abs(sto["t"] - stg["t"]) <= 1e-10 * sto["t"], (step, sto["t"], stg["t"])
        assert abs(sto["h"] - stg["h"]) <= 1e-9 * sto["h"], (step, sto["h"], stg["h"])
        fo = osim.get_state()
        fg = gsim.get_state(shape)
        for name, a, b in zip(("phi", "c", "q"), fg[:3], fo[:3]):
            e = _rel(a, b)
            assert e < 1e-10, (step, name, e, sto, stg)
        qs.add(stg["q"])
    assert qs == {1, 2}, qs
