"""GPU parity at the production tile shape: levels wider than one 64-cell tile
along x, as one patch and cut unevenly into patches, periodic and bounded,
every per-level kernel entry against the single-patch FP64 oracle.

The cuts are uneven (x at 71 of 136 in 3-D, x at 78 and y at 14 in 2-D), so
each patch holds two tiles along x and the patches meet inside a tile row;
on bounded levels the ghost corners at the cuts are zero-slope images of the
neighbour patch's cells, which the anisotropic / isotropic phase fluxes, the
isotropic EBS stencil and the orientation face gradients read.  The 2-D cuts
are even because the 2-D red-black sweep colours each patch from the parity
of jfirst + j (source/fortran/2d/ellipticfacops.m4:108-113), so an odd cut
changes the colouring there (in the reference too), while the 3-D colouring
ifirst + j + k is global.  Tolerances
as in test_gpu_parity.py: 1e-13 without transcendentals, 1e-12 with them,
bit-exact for tags.
"""
import ctypes as C

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu

SHAPES = {3: (136, 20, 12), 2: (150, 40, 1)}
# (name, cuts along x, cuts along y, periodic)
LAYOUTS = {
    3: [("one_periodic", [], [], True), ("two_periodic", [71], [], True),
        ("two_bounded", [71], [], False), ("four_bounded", [71], [9], False)],
    2: [("one_bounded", [], [], False), ("four_periodic", [78], [14], True),
        ("four_bounded", [78], [14], False)],
}
CASES = [(nd, lay) for nd in (3, 2) for lay in LAYOUTS[nd]]
IDS = [f"{nd}d-{lay[0]}" for nd, lay in CASES]


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Case:
    """Oracle patch over the whole domain and the device level cut into
    patches, with helpers moving cell and side fields between the two."""

    def __init__(self, ndim, layout, dev, p=None, g=1):
        from oracle import oracle as O
        from ampe_amd import hip as H
        self.O, self.H, self.dev = O, H, dev
        self.name, xc, yc, self.periodic = layout
        self.ndim = ndim
        nx, ny, nz = SHAPES[ndim]
        self.n = (nx, ny, nz)
        dx = tuple(p.dx) if p is not None else (1.0, 1.0, 1.0)
        self.pt = O.Patch(ndim, (nx, ny, nz), g=g, dx=dx)
        self.pt.G.periodic = 1 if self.periodic else 0
        xs, ys = [0] + xc + [nx], [0] + yc + [ny]
        self.boxes = []
        for j0, j1 in zip(ys[:-1], ys[1:]):
            for i0, i1 in zip(xs[:-1], xs[1:]):
                b = P.Box()
                b.lo[:] = [i0, j0, 0]
                b.hi[:] = [i1 - 1, j1 - 1, nz - 1 if ndim == 3 else 0]
                self.boxes.append(b)
        dom = P.Box()
        dom.lo[:] = [0, 0, 0]
        dom.hi[:] = [nx - 1, ny - 1, nz - 1 if ndim == 3 else 0]
        self.lev = H.Level(ndim, self.boxes, g, periodic=dom if self.periodic else None)
        self.g = g
        self.shape = (nz, ny, nx)

    # ---- cell fields
    def oracle_cells(self, vals):
        """ncomp x (nz, ny, nx) values -> oracle ghosted array, ghosts filled."""
        v = np.asarray(vals, dtype=np.float64).reshape((-1,) + self.shape)
        a = self.pt.cells(v.shape[0])
        self.pt.interior(a)[...] = v
        self.O.lib().o_ghost_fill(C.byref(self.pt.G), self.O.ptr(a), v.shape[0])
        return a

    def dev_cells(self, vals, fill=True):
        """The same values on the device level (ghosts filled by the device)."""
        import torch
        v = np.asarray(vals, dtype=np.float64).reshape((-1,) + self.shape)
        blocks = [self.lev.to_ghosted(x if self.ndim == 3 else x[0], ghost=self.g) for x in v]
        t = torch.from_numpy(np.concatenate(blocks)).to(self.dev)
        if fill:
            self.H.check(self.H.lib().ampe_ghost_fill(self.lev.h, _p(t), v.shape[0], None))
        return t

    def zeros_cells(self, ncomp=1):
        import torch
        return torch.zeros(ncomp * self.lev.ncg, dtype=torch.float64, device=self.dev)

    def interior_of_dev(self, t, ncomp=1):
        """Device cell field -> ncomp x (nz, ny, nx) interior values."""
        h = t.cpu().numpy().reshape(ncomp, self.lev.ncg)
        out = np.zeros((ncomp,) + self.shape)
        g, gz = self.g, (self.g if self.ndim == 3 else 0)
        for m in range(ncomp):
            for q, b in enumerate(self.boxes):
                blk = self.lev.ghost_view(h[m], q)
                blk = blk if self.ndim == 3 else blk[None]
                n = [b.hi[d] - b.lo[d] + 1 for d in range(3)]
                if self.ndim == 2:
                    n[2] = 1
                out[m, b.lo[2]:b.lo[2] + n[2], b.lo[1]:b.lo[1] + n[1], b.lo[0]:b.lo[0] + n[0]] = \
                    blk[gz:gz + n[2], g:g + n[1], g:g + n[0]]
        return out

    def interior_of_oracle(self, a, ncomp=1):
        return self.pt.interior(a).reshape((ncomp,) + self.shape)

    # ---- side fields
    def _face_blocks(self):
        """Per (patch, direction): (level offset, face shape (z, y, x), lower
        corner) of the patch's faces, and the oracle's block of direction d."""
        nx, ny, nz = self.n
        nzz = nz if self.ndim == 3 else 1
        o_off, o_shape = [], []
        off = 0
        for d in range(self.ndim):
            sh = [nzz + (d == 2), ny + (d == 1), nx + (d == 0)]
            o_off.append(off)
            o_shape.append(sh)
            off += int(np.prod(sh))
        blocks = []
        for q, b in enumerate(self.boxes):
            n = [b.hi[a] - b.lo[a] + 1 for a in range(3)]
            if self.ndim == 2:
                n[2] = 1
            for d in range(self.ndim):
                sh = (n[2] + (d == 2), n[1] + (d == 1), n[0] + (d == 0))
                # side_off is the patch's offset in the whole face array
                blocks.append((d, self.lev.side_off[q][d], sh,
                               (b.lo[2], b.lo[1], b.lo[0])))
        return o_off, o_shape, blocks

    def dev_sides(self, a_oracle):
        import torch
        o_off, o_shape, blocks = self._face_blocks()
        out = np.zeros(self.lev.nside_tot)
        for d, off, sh, lo in blocks:
            od = a_oracle[o_off[d]:o_off[d] + int(np.prod(o_shape[d]))].reshape(o_shape[d])
            out[off:off + int(np.prod(sh))] = od[lo[0]:lo[0] + sh[0], lo[1]:lo[1] + sh[1],
                                                 lo[2]:lo[2] + sh[2]].ravel()
        return torch.from_numpy(out).to(self.dev)

    def side_err(self, t, a_oracle):
        """Largest relative error of the level's faces (every patch's copy of a
        shared face) against the oracle's faces."""
        o_off, o_shape, blocks = self._face_blocks()
        h = t.cpu().numpy()
        got, ref = [], []
        for d, off, sh, lo in blocks:
            od = a_oracle[o_off[d]:o_off[d] + int(np.prod(o_shape[d]))].reshape(o_shape[d])
            ref.append(od[lo[0]:lo[0] + sh[0], lo[1]:lo[1] + sh[1], lo[2]:lo[2] + sh[2]].ravel())
            got.append(h[off:off + int(np.prod(sh))])
        return _rel(np.concatenate(got), np.concatenate(ref))


def _fields(shape, seed):
    rng = _rng(seed)
    phi = rng.uniform(0.0, 1.0, shape)
    qv = rng.normal(size=(4,) + shape)
    qv /= np.sqrt((qv ** 2).sum(0))
    c = rng.uniform(0.3, 0.5, shape)
    cl = rng.uniform(0.2, 0.6, shape)
    cs = rng.uniform(0.2, 0.6, shape)
    return phi, qv, c, cl, cs


def _phase_params(ndim, kind):
    if kind == "calphad":
        return P.polycrystal_params(ndim=ndim)
    p = P.two_grain_params(ndim=ndim)
    if kind == "iso":
        p.flux_type = P.FLUX_ISOTROPIC
    elif kind == "aniso":
        p.flux_type = P.FLUX_ANISOTROPIC
        p.aniso_nu = 0.05
        p.aniso_knumber = 4
    return p


@pytest.mark.parametrize("ndim,layout", CASES, ids=IDS)
def test_wide_phase_rhs(dev, ndim, layout):
    """Phase RHS for every flux form the dimension has, and the CALPHAD
    driving force, on wide uneven levels."""
    kinds = ["simple", "aniso", "calphad"] + (["iso"] if ndim == 2 else [])
    for kind in kinds:
        p = _phase_params(ndim, kind)
        cs_ = Case(ndim, layout, dev, p)
        O, H = cs_.O, cs_.H
        phi, q, c, cl, csv = _fields(cs_.shape, 11)
        of = [cs_.oracle_cells(v) for v in (phi, q, cl, csv)]
        r_o = cs_.pt.cells()
        O.lib().o_phase_rhs(C.byref(cs_.pt.G), C.byref(p), *[O.ptr(a) for a in of], O.ptr(r_o))
        df = [cs_.dev_cells(v) for v in (phi, q, cl, csv)]
        r = cs_.zeros_cells()
        H.check(H.lib().ampe_phase_rhs(cs_.lev.h, C.byref(p), *[_p(t) for t in df], _p(r), None))
        tol = 1e-13 if kind == "simple" else 1e-12
        e = _rel(cs_.interior_of_dev(r), cs_.interior_of_oracle(r_o))
        assert e < tol, (kind, e)


@pytest.mark.parametrize("ndim,layout", CASES, ids=IDS)
def test_wide_quat_grad_coefs_rhs(dev, ndim, layout):
    """Orientation gradients (cell modulus and face |grad q|^2), face
    coefficients, the orientation RHS from them and the projection."""
    p = _phase_params(ndim, "simple")
    cs_ = Case(ndim, layout, dev, p)
    O, H, pt = cs_.O, cs_.H, cs_.pt
    phi, q, c, cl, csv = _fields(cs_.shape, 13)
    phi_o, q_o = cs_.oracle_cells(phi), cs_.oracle_cells(q)
    gm_o, n2_o = pt.cells(), np.zeros(pt.nf)
    O.lib().o_quat_grad_modulus(C.byref(pt.G), C.byref(p), O.ptr(q_o), O.ptr(gm_o))
    O.lib().o_quat_face_norm2(C.byref(pt.G), C.byref(p), O.ptr(q_o), O.ptr(n2_o))
    phd, qd = cs_.dev_cells(phi), cs_.dev_cells(q)
    gm = cs_.zeros_cells()
    import torch
    n2 = torch.zeros(cs_.lev.nside_tot, dtype=torch.float64, device=dev)
    H.check(H.lib().ampe_quat_grad(cs_.lev.h, C.byref(p), _p(qd), _p(gm), _p(n2), None))
    assert _rel(cs_.interior_of_dev(gm), cs_.interior_of_oracle(gm_o)) < 1e-13
    assert cs_.side_err(n2, n2_o) < 1e-13
    # lagged norms differ from the current ones
    n2_o = n2_o * _rng(3).uniform(0.5, 1.5, n2_o.shape)
    n2d = cs_.dev_sides(n2_o)
    fc_o = np.zeros(pt.nf)
    O.lib().o_quat_face_coef(C.byref(pt.G), C.byref(p), O.ptr(phi_o), O.ptr(n2_o), O.ptr(fc_o))
    fc = torch.zeros(cs_.lev.nside_tot, dtype=torch.float64, device=dev)
    H.check(H.lib().ampe_quat_face_coefs(cs_.lev.h, C.byref(p), _p(phd), _p(n2d), _p(fc), None))
    assert cs_.side_err(fc, fc_o) < 1e-13
    r_o = pt.cells(4)
    O.lib().o_quat_rhs(C.byref(pt.G), C.byref(p), O.ptr(phi_o), O.ptr(q_o), O.ptr(n2_o),
                       O.ptr(r_o))
    r = cs_.zeros_cells(4)
    H.check(H.lib().ampe_quat_rhs_fc(cs_.lev.h, C.byref(p), _p(phd), _p(qd), _p(fc), _p(r),
                                     None))
    assert _rel(cs_.interior_of_dev(r, 4), cs_.interior_of_oracle(r_o, 4)) < 1e-12
    r1 = cs_.zeros_cells(4)
    H.check(H.lib().ampe_quat_rhs(cs_.lev.h, C.byref(p), _p(phd), _p(qd), _p(n2d), _p(r1), None))
    assert np.array_equal(cs_.interior_of_dev(r1, 4), cs_.interior_of_dev(r, 4))
    err = _rng(5).normal(size=(4,) + cs_.shape)
    corr_o, err_o = pt.cells(4), cs_.oracle_cells(err)
    O.lib().o_quat_project(C.byref(pt.G), C.byref(p), O.ptr(q_o), O.ptr(corr_o), O.ptr(err_o))
    corr, errd = cs_.zeros_cells(4), cs_.dev_cells(err)
    H.check(H.lib().ampe_quat_project(cs_.lev.h, C.byref(p), _p(qd), _p(corr), _p(errd), None))
    assert _rel(cs_.interior_of_dev(corr, 4), cs_.interior_of_oracle(corr_o, 4)) < 1e-13
    assert _rel(cs_.interior_of_dev(errd, 4), cs_.interior_of_oracle(err_o, 4)) < 1e-13


@pytest.mark.parametrize("ndim,layout", CASES, ids=IDS)
def test_wide_conc(dev, ndim, layout):
    """Composition: KKS-form coefficients and RHS, EBS coefficients with the
    regular and the isotropic stencil."""
    import torch
    p = _phase_params(ndim, "simple")
    p.diff_interp = b"p"
    cs_ = Case(ndim, layout, dev, p)
    O, H, pt = cs_.O, cs_.H, cs_.pt
    phi, q, c, cl, csv = _fields(cs_.shape, 17)
    ph_o, c_o, cl_o, cs_o = (cs_.oracle_cells(v) for v in (phi, c, cl, csv))
    phd, cd, cld, csd = (cs_.dev_cells(v) for v in (phi, c, cl, csv))
    Dc_o, Dp_o = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_conc_coeffs(C.byref(pt.G), C.byref(p), O.ptr(ph_o), O.ptr(cl_o), O.ptr(cs_o),
                          O.ptr(Dc_o), O.ptr(Dp_o))
    z = lambda: torch.zeros(cs_.lev.nside_tot, dtype=torch.float64, device=dev)  # noqa: E731
    Dc, Dp = z(), z()
    H.check(H.lib().ampe_conc_coeffs(cs_.lev.h, C.byref(p), _p(phd), _p(cld), _p(csd), _p(Dc),
                                     _p(Dp), None))
    assert cs_.side_err(Dc, Dc_o) < 1e-12
    assert cs_.side_err(Dp, Dp_o) < 1e-12
    r_o = pt.cells()
    O.lib().o_conc_rhs(C.byref(pt.G), C.byref(p), O.ptr(c_o), O.ptr(ph_o), O.ptr(Dc_o),
                       O.ptr(Dp_o), O.ptr(r_o))
    r = cs_.zeros_cells()
    H.check(H.lib().ampe_conc_rhs(cs_.lev.h, C.byref(p), _p(cd), _p(phd), _p(Dc), _p(Dp), _p(r),
                                  None))
    assert _rel(cs_.interior_of_dev(r), cs_.interior_of_oracle(r_o)) < 1e-12
    # EBS, regular stencil
    DL_o, DA_o = np.zeros(pt.nf), np.zeros(pt.nf)
    O.lib().o_ebs_coeffs(C.byref(pt.G), C.byref(p), O.ptr(ph_o), O.ptr(DL_o), O.ptr(DA_o))
    DL, DA = z(), z()
    H.check(H.lib().ampe_conc_coeffs_ebs(cs_.lev.h, C.byref(p), _p(phd), _p(DL), _p(DA), None))
    assert cs_.side_err(DL, DL_o) < 1e-13 and cs_.side_err(DA, DA_o) < 1e-13
    O.lib().o_conc_rhs_ebs(C.byref(pt.G), C.byref(p), O.ptr(cl_o), O.ptr(cs_o), O.ptr(DL_o),
                           O.ptr(DA_o), O.ptr(r_o))
    H.check(H.lib().ampe_conc_rhs_ebs(cs_.lev.h, C.byref(p), _p(cld), _p(csd), _p(DL), _p(DA),
                                      _p(r), None))
    assert _rel(cs_.interior_of_dev(r), cs_.interior_of_oracle(r_o)) < 1e-13
    # EBS, isotropic stencil (reads the corner ghosts)
    p.ebs_stencil = P.EBS_ISOTROPIC
    O.lib().o_conc_rhs_ebs_iso(C.byref(pt.G), C.byref(p), O.ptr(ph_o), O.ptr(cl_o), O.ptr(cs_o),
                               O.ptr(r_o))
    H.check(H.lib().ampe_conc_rhs_ebs_iso(cs_.lev.h, C.byref(p), _p(phd), _p(cld), _p(csd),
                                          _p(r), None))
    assert _rel(cs_.interior_of_dev(r), cs_.interior_of_oracle(r_o)) < 1e-13


@pytest.mark.parametrize("ndim,layout", CASES, ids=IDS)
def test_wide_kks_and_tags(dev, ndim, layout):
    """CALPHAD KKS over the ghosted boxes (interior compared; identical largest
    iteration count) and gradient tags, bit for bit."""
    import torch
    p = _phase_params(ndim, "calphad")
    cs_ = Case(ndim, layout, dev, p)
    O, H, pt = cs_.O, cs_.H, cs_.pt
    phi, q, c, cl, csv = _fields(cs_.shape, 19)
    ph_o, c_o = cs_.oracle_cells(phi), cs_.oracle_cells(c)
    cl_o, cs_o = pt.cells(), pt.cells()
    it_o = O.lib().o_kks_solve(C.byref(pt.G), C.byref(p), O.ptr(ph_o), O.ptr(c_o), O.ptr(c_o),
                               O.ptr(c_o), O.ptr(cl_o), O.ptr(cs_o))
    assert it_o >= 0
    phd, cd = cs_.dev_cells(phi), cs_.dev_cells(c)
    cld, csd = cs_.zeros_cells(), cs_.zeros_cells()
    it = C.c_int(0)
    rc = H.check(H.lib().ampe_kks_solve(cs_.lev.h, C.byref(p), _p(phd), _p(cd), _p(cd), _p(cd),
                                        _p(cld), _p(csd), C.byref(it), None))
    assert rc == 0 and it.value == it_o
    assert _rel(cs_.interior_of_dev(cld), cs_.interior_of_oracle(cl_o)) < 1e-12
    assert _rel(cs_.interior_of_dev(csd), cs_.interior_of_oracle(cs_o)) < 1e-12
    # tags
    N = int(np.prod(cs_.shape))
    tags_o = np.zeros(N, dtype=np.int32)
    O.lib().o_tag_cells(C.byref(pt.G), C.byref(p), O.ptr(ph_o), 0.3, O.iptr(tags_o))
    tags = torch.zeros(N, dtype=torch.int32, device=dev)
    H.check(H.lib().ampe_tag_cells(cs_.lev.h, C.byref(p), _p(phd), 0.3, _p(tags), None))
    got = cs_.lev.from_interior(tags.cpu().numpy())
    ref = tags_o.reshape(cs_.shape if ndim == 3 else cs_.shape[1:])
    assert np.array_equal(got, ref)
    assert 0 < tags_o.sum() < N


@pytest.mark.parametrize("ndim,layout", CASES, ids=IDS)
def test_wide_fac_smooth_residual(dev, ndim, layout):
    """FP64 red-black sweeps and residual of the phase block (C field, constant
    D) and of the composition block (D on faces): the colouring follows the
    global index, so uneven cuts must not change a value."""
    p = _phase_params(ndim, "simple")
    cs_ = Case(ndim, layout, dev, p)
    O, H, pt = cs_.O, cs_.H, cs_.pt
    phi, q, c, cl, csv = _fields(cs_.shape, 29)
    gamma = 1.0e-3
    ph_o = cs_.oracle_cells(phi)
    C_o = pt.cells()
    O.lib().o_phase_fac_C(C.byref(pt.G), C.byref(p), gamma, O.ptr(ph_o), None, O.ptr(C_o))
    phd = cs_.dev_cells(phi)
    Cd = cs_.zeros_cells()
    H.check(H.lib().ampe_fac_phase_coeffs(cs_.lev.h, C.byref(p), gamma, _p(phd), None, _p(Cd),
                                          None))
    assert _rel(cs_.interior_of_dev(Cd), cs_.interior_of_oracle(C_o)) < 1e-14
    rhs_v = _rng(31).normal(size=cs_.shape)
    rhs_o, rhsd = cs_.oracle_cells(rhs_v), cs_.dev_cells(rhs_v)
    Dside = -gamma * _rng(37).uniform(1.0, 5.0, pt.nf)
    Dd = cs_.dev_sides(Dside)
    for variant in ("phase", "conc"):
        op_o, op_h = O.EllipticOp(), H.EllipticOp()
        if variant == "phase":
            op_o.D = op_h.D = None
            op_o.D_const = op_h.D_const = -gamma * p.epsilon_phi ** 2
            op_o.C, op_h.C = C_o.ctypes.data, Cd.data_ptr()
            op_o.M_const = op_h.M_const = p.phi_mobility
        else:
            op_o.D, op_h.D = Dside.ctypes.data, Dd.data_ptr()
            op_o.C = op_h.C = None
            op_o.C_const = op_h.C_const = 1.0
            op_o.M_const = op_h.M_const = 1.0
        op_o.M = op_h.M = None
        op_o.dx = op_h.dx = (C.c_double * 3)(*p.dx)
        u_o = pt.cells()
        flux = np.zeros(pt.nf)
        m_o = O.lib().o_fac_smooth(C.byref(pt.G), C.byref(op_o), O.ptr(rhs_o), O.ptr(u_o), 3,
                                   -1.0, O.ptr(flux))
        ud = cs_.zeros_cells()
        mres = C.c_double(0)
        H.check(H.lib().ampe_fac_smooth(cs_.lev.h, C.byref(op_h), _p(rhsd), _p(ud), 3, -1.0,
                                        C.byref(mres), None))
        assert _rel(cs_.interior_of_dev(ud), cs_.interior_of_oracle(u_o)) < 1e-13, variant
        assert abs(mres.value - m_o) <= 1e-13 * abs(m_o), variant
        res_o = pt.cells()
        rms_o = O.lib().o_fac_residual(C.byref(pt.G), C.byref(op_o), O.ptr(rhs_o), O.ptr(u_o),
                                       O.ptr(res_o), O.ptr(flux))
        resd = cs_.zeros_cells()
        rms = C.c_double(0)
        H.check(H.lib().ampe_fac_residual(cs_.lev.h, C.byref(op_h), _p(rhsd), _p(ud), _p(resd),
                                          C.byref(rms), None))
        assert _rel(cs_.interior_of_dev(resd), cs_.interior_of_oracle(res_o)) < 1e-12, variant
        assert abs(rms.value - rms_o) <= 1e-12 * rms_o, variant


@pytest.mark.parametrize("ndim,layout", CASES, ids=IDS)
def test_wide_reductions(dev, ndim, layout):
    """Dot products and WRMS norms over uneven patches of six-component
    fields, against the oracle (fixed-order sums in both)."""
    cs_ = Case(ndim, layout, dev)
    O, H, pt = cs_.O, cs_.H, cs_.pt
    rng = _rng(41)
    x, y = rng.normal(size=(6,) + cs_.shape), rng.normal(size=(6,) + cs_.shape)
    w = rng.uniform(0.5, 2.0, size=(6,) + cs_.shape)
    xo, yo, wo = (cs_.oracle_cells(v) for v in (x, y, w))
    d_o = O.lib().o_dot(C.byref(pt.G), 6, O.ptr(xo), O.ptr(yo))
    w_o = O.lib().o_wrms(C.byref(pt.G), 6, O.ptr(xo), O.ptr(wo))
    xd, yd, wd = (cs_.dev_cells(v) for v in (x, y, w))
    out = C.c_double(0)
    H.check(H.lib().ampe_vec_dot(cs_.lev.h, 6, _p(xd), _p(yd), C.byref(out), None))
    assert abs(out.value - d_o) <= 1e-12 * np.abs(x * y).sum()
    H.check(H.lib().ampe_vec_wrms(cs_.lev.h, 6, _p(xd), _p(wd), C.byref(out), None))
    assert abs(out.value - w_o) <= 1e-13 * w_o
