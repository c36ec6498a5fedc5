"""CPU checks pinning the oracle's anisotropic phase fluxes, orientation
face coefficients, orientation mobility, orientation RHS with its Lagrange
projection, the CVODE projection of q, gradient tagging and the vector
reductions against mathematics that the oracle's code does not restate:

* anisotropic fluxes (source/fortran/2d/quatrhs.m4:154-260,
  3d/quatrhs.m4:149-356): on linear phase fields the stencil gradients are
  exact, so every face flux must equal the continuum flux of the anisotropic
  gradient energy, J = eps(theta)^2 grad phi + eps(theta) eps'(theta) perp(grad phi)
  in 2-D (Kobayashi form) and J = gamma^2 g + gamma |g| d gamma / d n in 3-D with
  the cubic gamma(n) = eps (1 - 3 eps4) (1 + 4 eps4 / (1 - 3 eps4) sum m_i^4),
  m = R(q) n; the rotation R(q) comes from scipy, not from the oracle's
  quaternion products.  nu = 0 / eps4 = 0 reduce to the simple flux, and a
  reflection of phi across the diagonal reflects the 2-D flux (4-fold, psi = 0);
* orientation face coefficient (2d/quatfacops.m4:14-122): fc = -(1/|g|) de/d|g|
  of the orientation energy density e = 2 H T p1 |g| + eps_q^2 / 2 p2 |g|^2,
  differentiated numerically;
* orientation RHS (QuatFACOps::evaluateRHS + add_quat_proj_op): tangent to the
  sphere (q . rhs = 0 per cell), zero for a uniform twist (a harmonic map),
  second-order convergence to -M_q fc theta'' q_perp for q = (cos theta(x),
  sin theta(x), 0, 0), and equivariance under a constant 4x4 rotation of q;
* projection (2d/quatfacops.m4:940-997): |q + corr| = 1 with corr parallel to
  q, and err loses exactly its component along q;
* tagging (2d/quattag.m4:9-55): numpy central differences on random fields,
  and the >= threshold at an exactly representable equality;
* dot / WRMS (Sundials_SAMRAIVector): numpy sums over interior cells only.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from ampe_amd import params as P


@pytest.fixture(scope="module")
def O():
    from oracle import oracle as O
    O.lib()
    return O


def _grid(pt):
    """Cell-centre coordinates (x, y, z) of every ghosted cell, shape gshape."""
    g, gz = pt.g, pt.gz
    kk, jj, ii = np.meshgrid(np.arange(pt.nz + 2 * gz) - gz, np.arange(pt.ny + 2 * g) - g,
                             np.arange(pt.nx + 2 * g) - g, indexing="ij")
    dx = pt.G.dx
    return ii * dx[0], jj * dx[1], kk * dx[2]


def _faces(pt, arr):
    """Face arrays per direction, shaped (nz[+1], ny[+1], nx[+1])."""
    nz, ny, nx = pt.nz, pt.ny, pt.nx
    out, off = [], 0
    for d in range(pt.ndim):
        sh = (nz + (d == 2), ny + (d == 1), nx + (d == 0))
        n = int(np.prod(sh))
        out.append(arr[off:off + n].reshape(sh))
        off += n
    return out


def _flux(O, pt, p, phi, q):
    fl = np.zeros(pt.nf)
    O.lib().o_phase_flux(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(q), O.ptr(fl))
    return _faces(pt, fl)


def _aniso_params(ndim, eps=1.3, nu=0.05, k=4):
    p = P.two_grain_params(ndim=ndim)
    p.flux_type = P.FLUX_ANISOTROPIC
    p.epsilon_phi = eps
    p.aniso_nu = nu
    p.aniso_knumber = k
    p.qlen = 4
    return p


# ------------------------------------------------------------ 2-D anisotropy
@pytest.mark.parametrize("a,b", [(0.7, 0.0), (0.0, -0.4), (0.5, 0.3), (-0.2, 0.6)])
@pytest.mark.parametrize("psi", [0.0, 0.3, 1.1])
def test_aniso_2d_linear_field_is_continuum_flux(O, a, b, psi):
    """phi = a x + b y, uniform q0 = cos(psi / 2): every x and y face carries
    J = eps(theta)^2 (a, b) + eps eps' (-b, a), theta = atan(b / a) (pi / 2 when
    a = 0), eps(theta) = eps (1 + nu cos k(theta - psi))."""
    eps, nu, k = 1.3, 0.05, 4
    p = _aniso_params(2, eps, nu, k)
    pt = O.Patch(2, (9, 7, 1), dx=(0.5, 0.25, 1.0))
    x, y, _ = _grid(pt)
    phi = (a * x + b * y)[None]
    q = pt.cells(4)
    q[0] = np.cos(psi / 2)
    q[1] = np.sin(psi / 2)
    fx, fy = _flux(O, pt, p, phi, q)
    theta = np.arctan(b / a) if a != 0 else 0.5 * np.pi
    et = eps * (1 + nu * np.cos(k * (theta - psi)))
    de = -eps * nu * k * np.sin(k * (theta - psi))
    jx = et * et * a - et * de * b
    jy = et * et * b + et * de * a
    scale = eps * eps * np.hypot(a, b)
    assert np.allclose(fx, jx, rtol=0, atol=1e-13 * scale)
    assert np.allclose(fy, jy, rtol=0, atol=1e-13 * scale)


def test_aniso_2d_front_along_x_closed_form(O):
    """Planar front along x with uniform orientation psi = 2 acos q0: theta = 0,
    so the normal flux is eps^2 (1 + nu cos k psi)^2 dphi/dx and the transverse
    (y-face) flux eps^2 (1 + nu cos k psi) k nu sin(k psi) dphi/dx."""
    eps, nu, k, psi = 0.9, 0.05, 4, 0.5
    p = _aniso_params(2, eps, nu, k)
    pt = O.Patch(2, (6, 5, 1))
    x, y, _ = _grid(pt)
    phi = (0.25 * x)[None]
    q = pt.cells(4)
    q[0] = np.cos(psi / 2)
    q[3] = np.sin(psi / 2)
    fx, fy = _flux(O, pt, p, phi, q)
    assert np.allclose(fx, eps ** 2 * (1 + nu * np.cos(k * psi)) ** 2 * 0.25, rtol=1e-14)
    assert np.allclose(fy, eps ** 2 * (1 + nu * np.cos(k * psi)) * k * nu * np.sin(k * psi) * 0.25,
                       rtol=1e-13)


def test_aniso_2d_nu_zero_is_simple_flux(O):
    rng = np.random.Generator(np.random.Philox(3))
    pt = O.Patch(2, (11, 8, 1), dx=(0.3, 0.2, 1.0))
    phi = pt.load(pt.cells(), rng.uniform(0, 1, (1, 8, 11)))
    q = pt.load(pt.cells(4), rng.normal(size=(4, 1, 8, 11)))
    pa = _aniso_params(2, nu=0.0)
    ps = P.two_grain_params(ndim=2, flux_type=P.FLUX_SIMPLE)
    ps.epsilon_phi = pa.epsilon_phi
    for fa, fs in zip(_flux(O, pt, pa, phi, q), _flux(O, pt, ps, phi, None)):
        assert np.allclose(fa, fs, rtol=1e-14, atol=1e-14 * np.abs(fs).max())


def test_aniso_2d_reflection_symmetry(O):
    """psi = 0, k = 4: eps(pi/2 - theta) = eps(theta), so reflecting phi across
    the diagonal reflects the flux: x faces of phi^T are y faces of phi."""
    n = 14
    rng = np.random.Generator(np.random.Philox(5))
    f = rng.uniform(0, 1, (n, n))
    p = _aniso_params(2, eps=1.1, nu=0.2, k=4)
    out = []
    for field in (f, f.T.copy()):
        pt = O.Patch(2, (n, n, 1))
        phi = pt.load(pt.cells(), field)
        q = pt.cells(4)
        q[0] = 1.0
        out.append(_flux(O, pt, p, phi, q))
    (fx, fy), (gx, gy) = out
    scale = np.abs(fx).max()
    assert np.allclose(gx[0], fy[0].T, rtol=0, atol=1e-13 * scale)
    assert np.allclose(gy[0], fx[0].T, rtol=0, atol=1e-13 * scale)


# ------------------------------------------------------------ 3-D anisotropy
def _cubic_flux(eps, eps4, g, q):
    """Continuum flux of the cubic gradient energy 1/2 gamma(n)^2 |g|^2 for a
    crystal rotated by the unit quaternion q = (w, x, y, z)."""
    R = Rotation.from_quat([q[1], q[2], q[3], q[0]])
    gn = np.linalg.norm(g)
    n = g / gn
    m = R.apply(n)
    s4 = np.sum(m ** 4)
    gam = eps * (1 - 3 * eps4) * (1 + 4 * eps4 / (1 - 3 * eps4) * s4)
    dgam_m = 16 * eps * eps4 * (m ** 3 - m * s4)  # tangential d gamma / d m
    return gam * gam * g + gam * gn * R.inv().apply(dgam_m)


@pytest.mark.parametrize("gdir", [(1, 0, 0), (0, 0, 1), (1, 1, 0), (0.3, -0.8, 0.5),
                                  (1, 1, 1), (-0.2, 0.1, 0.9)])
@pytest.mark.parametrize("qi", [0, 1, 2])
def test_aniso_3d_linear_field_is_cubic_continuum_flux(O, gdir, qi):
    eps, eps4 = 1.2, 0.05
    qs = [np.array([1.0, 0, 0, 0]),
          np.array([np.cos(np.pi / 8), 0, 0, np.sin(np.pi / 8)]),
          np.array([0.8, -0.1, 0.5, 0.3])]
    qv = qs[qi] / np.linalg.norm(qs[qi])
    g = 0.6 * np.asarray(gdir, dtype=float)
    p = _aniso_params(3, eps, eps4)
    pt = O.Patch(3, (5, 4, 6), dx=(0.5, 0.4, 0.25))
    x, y, z = _grid(pt)
    phi = (g[0] * x + g[1] * y + g[2] * z)[None]
    q = pt.cells(4)
    for m in range(4):
        q[m] = qv[m]
    J = _cubic_flux(eps, eps4, g, qv)
    scale = eps * eps * np.linalg.norm(g)
    for d, f in enumerate(_flux(O, pt, p, phi, q)):
        assert np.allclose(f, J[d], rtol=0, atol=1e-12 * scale), (d, f.ravel()[:3], J)


def test_aniso_3d_axis_front_closed_form(O):
    """q = identity, front along an axis: eps^2 (1 + eps4)^2 dphi/dx_d, and no
    transverse flux."""
    eps, eps4 = 0.7, 0.04
    p = _aniso_params(3, eps, eps4)
    pt = O.Patch(3, (4, 5, 3))
    x, y, z = _grid(pt)
    for d, coord in enumerate((x, y, z)):
        phi = (0.5 * coord)[None]
        q = pt.cells(4)
        q[0] = 1.0
        fl = _flux(O, pt, p, phi, q)
        for e in range(3):
            expect = eps ** 2 * (1 + eps4) ** 2 * 0.5 if e == d else 0.0
            assert np.allclose(fl[e], expect, rtol=1e-13, atol=1e-15)


def test_aniso_3d_eps4_zero_is_simple_flux(O):
    rng = np.random.Generator(np.random.Philox(7))
    pt = O.Patch(3, (7, 6, 5), dx=(0.3, 0.2, 0.1))
    phi = pt.load(pt.cells(), rng.uniform(0, 1, (5, 6, 7)))
    qv = rng.normal(size=(4, 5, 6, 7))
    q = pt.load(pt.cells(4), qv / np.sqrt((qv ** 2).sum(0)))
    pa = _aniso_params(3, nu=0.0)
    ps = P.two_grain_params(ndim=3, flux_type=P.FLUX_SIMPLE)
    ps.epsilon_phi = pa.epsilon_phi
    for fa, fs in zip(_flux(O, pt, pa, phi, q), _flux(O, pt, ps, phi, None)):
        assert np.allclose(fa, fs, rtol=1e-14, atol=1e-14 * np.abs(fs).max())


# ------------------------------------------------------- orientation terms
def _quat_params(ndim=2):
    p = P.two_grain_params(ndim=ndim)
    p.floor_type = b"m"
    p.gradient_floor = 1e-3
    p.avg_type = b"a"
    return p


@pytest.mark.parametrize("phi_l,phi_r", [(1.0, 1.0), (0.2, 0.6), (0.0, 0.0), (0.9, 0.3)])
@pytest.mark.parametrize("g", [0.05, 0.7, 3.0])
def test_face_coef_is_energy_derivative(O, phi_l, phi_r, g):
    """fc = -(1/|g|) de/d|g| with e(|g|) = 2 H T p1(phi_f) |g| + eps_q^2 / 2 p2(phi_f)
    |g|^2 at the face average phi_f, the derivative by central differences."""
    p = _quat_params()
    pt = O.Patch(2, (2, 1, 1))
    phi = pt.cells()
    phi[0, 0, :, :2] = phi_l  # cells -1, 0 (x face 0 between them... and 1)
    phi[0, 0, :, 2:] = phi_r
    n2 = np.full(pt.nf, g * g)
    fc = np.zeros(pt.nf)
    O.lib().o_quat_face_coef(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(n2), O.ptr(fc))
    # x face 1 lies between cell 0 (phi_l) and cell 1 (phi_r)
    pf = 0.5 * (phi_l + phi_r)
    p1 = O.lib().o_interp_func(pf, p.orient_interp1)
    p2 = O.lib().o_interp_func(pf, p.orient_interp2)

    def e(s):
        return 2 * p.H_parameter * p.temperature * p1 * s + 0.5 * p.epsilon_q ** 2 * p2 * s * s

    h = 1e-6 * g
    dedg = (e(g + h) - e(g - h)) / (2 * h)
    expect = -dedg / g
    got = _faces(pt, fc)[0][0, 0, 1]
    assert abs(got - expect) <= 1e-7 * abs(expect) + 1e-12, (got, expect)


def test_quat_mobility_closed_form(O):
    """'p' mobility: M_q(0) = M_max, M_q(1) = M_min, M_q(1/2) = the mean (pbg
    is 1/2 there), decreasing; 'e': the same end points."""
    p = _quat_params()
    p.quat_mobility, p.min_quat_mobility = 200.0, 2.0
    pt = O.Patch(2, (5, 1, 1))
    for t in (b"p", b"e"):
        p.quat_mobility_type = t
        p.quat_mobility_scale = 3.0
        phi = pt.cells()
        phi[0, 0, 1, 1:6] = [0.0, 0.25, 0.5, 0.75, 1.0]
        mq = pt.cells()
        O.lib().o_quat_mobility(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(mq))
        m = mq[0, 0, 1, 1:6]
        assert m[0] == pytest.approx(200.0, rel=1e-15)
        assert m[-1] == pytest.approx(2.0, rel=1e-15)
        assert np.all(np.diff(m) < 0)
        if t == b"p":
            assert m[2] == pytest.approx(101.0, rel=1e-15)


def _quat_rhs(O, pt, p, phi, q, n2):
    r = pt.cells(4)
    O.lib().o_quat_rhs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(q), O.ptr(n2), O.ptr(r))
    return r


@pytest.mark.parametrize("ndim", [2, 3])
def test_quat_rhs_tangent_and_rotation_equivariant(O, ndim):
    """q . rhs = 0 in every cell (the Lagrange term keeps |q| = 1 to first
    order), and rhs(R q) = R rhs(q) for a constant orthogonal 4x4 R."""
    rng = np.random.Generator(np.random.Philox(11))
    n = (9, 8, 7)
    pt = O.Patch(ndim, n, dx=(0.5, 0.5, 0.5))
    shape = (pt.nz, pt.ny, pt.nx)
    p = _quat_params(ndim)
    phi = pt.load(pt.cells(), rng.uniform(0, 1, shape))
    qv = rng.normal(size=(4,) + shape)
    qv /= np.sqrt((qv ** 2).sum(0))
    q = pt.load(pt.cells(4), qv)
    n2 = rng.uniform(0.1, 4.0, pt.nf)
    r = _quat_rhs(O, pt, p, phi, q, n2)
    dotp = (pt.interior(r) * pt.interior(q)).sum(0)
    assert np.abs(dotp).max() < 1e-12 * np.abs(pt.interior(r)).max()
    assert np.abs(pt.interior(r)).max() > 0
    Rm, _ = np.linalg.qr(rng.normal(size=(4, 4)))
    qr = np.einsum("mn,n...->m...", Rm, q)
    r2 = _quat_rhs(O, pt, p, phi, qr, n2)
    rr = np.einsum("mn,n...->m...", Rm, r)
    assert np.allclose(pt.interior(r2), pt.interior(rr), rtol=0,
                       atol=1e-12 * np.abs(pt.interior(r)).max())


def test_quat_rhs_uniform_twist_is_steady(O):
    """q = (cos a x, sin a x, 0, 0) on uniform phi with uniform face norms: a
    constant-rate twist is a harmonic map, the projected RHS vanishes (the
    discrete Laplacian of q is parallel to q)."""
    n = 16
    pt = O.Patch(2, (n, 4, 1))
    p = _quat_params()
    x, _, _ = _grid(pt)
    a = 2 * np.pi / n
    q = pt.cells(4)
    q[0], q[1] = np.cos(a * x), np.sin(a * x)
    phi = pt.cells() + 0.7
    n2 = np.full(pt.nf, a * a)
    r = _quat_rhs(O, pt, p, phi, q, n2)
    fc, mq = np.zeros(pt.nf), pt.cells()
    O.lib().o_quat_face_coef(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(n2), O.ptr(fc))
    O.lib().o_quat_mobility(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(mq))
    # size of either term before the projection cancels them
    scale = mq.max() * np.abs(fc).max() * (2 - 2 * np.cos(a))
    assert np.abs(pt.interior(r)).max() < 1e-13 * scale


def test_quat_rhs_converges_to_projected_continuum(O):
    """q = (cos th(x), sin th(x), 0, 0), th = b x^2, uniform phi and face norms
    (so uniform fc < 0): rhs -> -M_q fc th''(x) q_perp, q_perp = (-sin th, cos th),
    with second-order error."""
    p = _quat_params()
    errs = []
    for n in (32, 64):
        L = 1.0
        h = L / n
        pt = O.Patch(2, (n, 3, 1), dx=(h, h, 1.0))
        x, _, _ = _grid(pt)
        b = 0.8
        th = b * x * x
        q = pt.cells(4)
        q[0], q[1] = np.cos(th), np.sin(th)
        phi = pt.cells() + 1.0
        n2 = np.full(pt.nf, 2.0)
        fcv = np.zeros(pt.nf)
        O.lib().o_quat_face_coef(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(n2), O.ptr(fcv))
        fc = fcv[0]
        assert fc < 0 and np.allclose(fcv, fc)
        mq = pt.cells()
        O.lib().o_quat_mobility(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(mq))
        M = mq.ravel()[0]
        r = _quat_rhs(O, pt, p, phi, q, n2)
        thi = pt.interior(th[None])[0]
        expect = -M * fc * 2 * b * np.stack([-np.sin(thi), np.cos(thi)])
        got = pt.interior(r)[:2]
        errs.append(np.abs(got - expect).max() / np.abs(expect).max())
        assert np.abs(pt.interior(r)[2:]).max() == 0.0
    assert errs[1] < 0.3 * errs[0] and errs[1] < 1e-2, errs


def test_quat_project_closed_form(O):
    rng = np.random.Generator(np.random.Philox(13))
    pt = O.Patch(3, (5, 4, 3))
    shape = (3, 4, 5)
    p = _quat_params(3)
    q = pt.load(pt.cells(4), rng.normal(size=(4,) + shape) * 1.3)
    err0 = pt.load(pt.cells(4), rng.normal(size=(4,) + shape))
    corr, err = pt.cells(4), err0.copy()
    O.lib().o_quat_project(C.byref(pt.G), C.byref(p), O.ptr(q), O.ptr(corr), O.ptr(err))
    qi, ci = pt.interior(q), pt.interior(corr)
    ei, e0 = pt.interior(err), pt.interior(err0)
    nq = np.sqrt((qi ** 2).sum(0))
    assert np.allclose(np.sqrt(((qi + ci) ** 2).sum(0)), 1.0, rtol=1e-14)
    assert np.allclose(ci, qi * (1 / nq - 1), rtol=1e-13, atol=1e-15)
    qh = qi / nq
    along = (e0 * qh).sum(0)
    assert np.abs((ei * qh).sum(0)).max() < 1e-14 * np.abs(e0).max()
    assert np.allclose(ei + along * qh, e0, rtol=0, atol=1e-14 * np.abs(e0).max())


# ------------------------------------------------------------------ tagging
@pytest.mark.parametrize("ndim", [2, 3])
def test_tags_match_numpy_central_differences(O, ndim):
    rng = np.random.Generator(np.random.Philox(17))
    n = (7, 6, 5)
    dx = (0.5, 0.25, 2.0)
    pt = O.Patch(ndim, n, dx=dx)
    shape = (pt.nz, pt.ny, pt.nx)
    f = rng.uniform(0, 1, shape)
    phi = pt.load(pt.cells(), f)
    tol = 0.35
    tags = np.zeros(pt.N, dtype=np.int32)
    p = P.two_grain_params(ndim=ndim)
    O.lib().o_tag_cells(C.byref(pt.G), C.byref(p), O.ptr(phi), tol, O.iptr(tags))
    g2 = np.zeros(shape)
    for d in range(ndim):
        ax = 2 - d
        grad = (np.roll(f, -1, axis=ax) - np.roll(f, 1, axis=ax)) / (2 * dx[d])
        g2 += (grad * dx[d]) ** 2
    expect = (g2 >= tol * tol).astype(np.int32).ravel()
    assert np.array_equal(tags, expect)
    assert 0 < tags.sum() < pt.N


def test_tag_threshold_is_inclusive(O):
    """phi = x / 4 with dx = 1: |grad phi dx| = 1/4 exactly, tagged at tol = 1/4
    (>=), untagged just above."""
    pt = O.Patch(2, (6, 4, 1))
    x, _, _ = _grid(pt)
    phi = (0.25 * x)[None]
    p = P.two_grain_params()
    for tol, want in ((0.25, 1), (0.25 * (1 + 1e-12), 0)):
        tags = np.zeros(pt.N, dtype=np.int32)
        O.lib().o_tag_cells(C.byref(pt.G), C.byref(p), O.ptr(phi), tol, O.iptr(tags))
        assert np.all(tags == want)


# --------------------------------------------------------------- reductions
@pytest.mark.parametrize("ndim", [2, 3])
def test_dot_and_wrms_are_interior_sums(O, ndim):
    """dot = cell volume * sum x y, WRMS = sqrt(mean (x w)^2) over interior
    cells of every component; ghost values (set to garbage) do not enter."""
    rng = np.random.Generator(np.random.Philox(19))
    dx = (0.5, 0.25, 2.0)
    pt = O.Patch(ndim, (6, 5, 4), dx=dx)
    x = rng.normal(size=pt.cells(3).shape) * 1e3
    y = rng.normal(size=pt.cells(3).shape) * 1e3
    w = rng.uniform(0.5, 2.0, pt.cells(3).shape) * 1e3
    for a in (x, y, w):
        pt.interior(a)[...] = rng.normal(size=pt.interior(a).shape)
    cv = dx[0] * dx[1] * (dx[2] if ndim == 3 else 1.0)
    d = O.lib().o_dot(C.byref(pt.G), 3, O.ptr(x), O.ptr(y))
    assert d == pytest.approx(cv * np.sum(pt.interior(x) * pt.interior(y)), rel=1e-13)
    wr = O.lib().o_wrms(C.byref(pt.G), 3, O.ptr(x), O.ptr(w))
    assert wr == pytest.approx(np.sqrt(np.mean((pt.interior(x) * pt.interior(w)) ** 2)),
                               rel=1e-13)


# ------------------------------------------------------------ phase RHS
@pytest.mark.parametrize("model", ["quadratic", "calphad"])
@pytest.mark.parametrize("orient", [False, True])
@pytest.mark.parametrize("ph", [0.2, 0.55, 0.9])
def test_phase_rhs_is_variational_derivative(O, model, orient, ph):
    """Uniform phi (no flux divergence), q linear in x (uniform |grad q|): the
    phase RHS is -M dE/dphi of the cell's energy
    E(phi) = w g(phi) + 2 H T p1(phi) |grad q| + eps_q^2 / 2 p2(phi) |grad q|^2
             + (1 - h) f_L(c_L) + h f_S(c_S),
    with (c_L, c_S) the KKS phase concentrations of the fixed c at each phi, so
    the driving force h'[(f_L - f_S) - mu (c_L - c_S)] is the derivative taken
    along the KKS constraint; dE/dphi by central differences."""
    if model == "quadratic":
        p = P.two_grain_params(ndim=2)
        p.quad_ceq[0], p.quad_ceq[1] = 0.40, 0.45
        c = 0.43
    else:
        p = P.polycrystal_params(ndim=2)
        c = 0.41
    p.vm_L = p.vm_A = 1.0e-6  # energies in J/mol per unit volume factor 1
    p.conc_interp = p.energy_interp
    p.evolve_quat = 1 if orient else 0
    p.with_conc = 1
    p.flux_type = P.FLUX_SIMPLE
    pt = O.Patch(2, (3, 3, 1), dx=(0.5, 0.5, 1.0))
    x, _, _ = _grid(pt)
    alpha = np.array([0.3, -0.2, 0.1, 0.25])
    q = pt.cells(4)
    for m in range(4):
        q[m] = alpha[m] * x
    gm = np.sqrt(np.sum(alpha ** 2))  # |grad q| of the linear field

    def kks(phv):
        cl, cs = C.c_double(c), C.c_double(c)
        assert O.lib().o_kks_cell(C.byref(p), c, phv, C.byref(cl), C.byref(cs)) >= 0
        return cl.value, cs.value

    def energy(phv):
        cl, cs = kks(phv)
        h = O.lib().o_interp_func(phv, p.energy_interp)
        e = p.phi_well_scale * O.lib().o_well_func(phv, p.phi_well_type)
        e += (1 - h) * O.lib().o_fenergy(C.byref(p), cl, 0) + h * O.lib().o_fenergy(C.byref(p), cs, 1)
        if orient:
            e += 2 * p.H_parameter * p.temperature * O.lib().o_interp_func(phv, p.orient_interp1) * gm
            e += 0.5 * p.epsilon_q ** 2 * O.lib().o_interp_func(phv, p.orient_interp2) * gm * gm
        return e

    cl0, cs0 = kks(ph)
    phi = pt.cells() + ph
    clf, csf = pt.cells() + cl0, pt.cells() + cs0
    r = pt.cells()
    O.lib().o_phase_rhs(C.byref(pt.G), C.byref(p), O.ptr(phi), O.ptr(q), O.ptr(clf),
                        O.ptr(csf), O.ptr(r))
    hstep = 1e-5
    dE = (energy(ph + hstep) - energy(ph - hstep)) / (2 * hstep)
    expect = -p.phi_mobility * dE
    got = pt.interior(r)
    assert np.allclose(got, expect, rtol=2e-6, atol=0), (got.ravel()[0], expect)
