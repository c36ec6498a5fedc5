"""Oracle restatement of the orientation-field transfers (TEST
INFRASTRUCTURE ONLY: used by tests/ to check ampe_quat_coarsen /
ampe_quat_refine) on a periodic domain, ratio 2.

quatcoarsen (source/fortran/2d/quatcoarsen.m4, 3d/quatcoarsen.m4, called by
QuatWeightedAverage::coarsen, source/QuatWeightedAverage.cc:100-171): the sum
of q dVf over the children (k, then j, then i fastest) times 1/dVc, then
quatnorm (source/fortran/quat.f:1043-1083, q/|q| with |q| summed over the
components in order, 0 when |q| < 1e-15).  quatlinrefine
(source/fortran/2d/quatlinrefine.m4:13-114, 3d/quatlinrefine.m4, called by
QuatLinearRefine::refine, source/QuatLinearRefine.cc:107-174): the
(bi/tri)linear form with x = ((r + 1/2) dxf - dxc/2) / dxc, shifted down one
coarse cell when negative, then quatnorm.  The symmetry rotation with index 0
is the identity (the reference's qlen 1 / 2 behaviour)."""
import numpy as np


def _norm(q):
    s = np.zeros(q.shape[1:])
    for m in range(q.shape[0]):
        s = s + q[m] * q[m]
    mg = np.sqrt(s)
    minv = np.where(mg < 1.0e-15, 0.0, 1.0 / np.where(mg < 1.0e-15, 1.0, mg))
    return q * minv


def coarsen(qf, dxf):
    """qf: (ql, [nz,] ny, nx) -> (ql, [nz/2,] ny/2, nx/2)."""
    ndim = qf.ndim - 1
    dVf = dxf[0] * dxf[1]
    dVc = (2.0 * dxf[0]) * (2.0 * dxf[1])
    if ndim == 3:
        dVf = dVf * dxf[2]
        dVc = dVc * (2.0 * dxf[2])
    dVcinv = 1.0 / dVc
    cshape = (qf.shape[0],) + tuple(s // 2 for s in qf.shape[1:])
    acc = np.zeros(cshape)
    for r2 in range(2 if ndim == 3 else 1):
        for r1 in range(2):
            for r0 in range(2):
                if ndim == 3:
                    acc = acc + qf[:, r2::2, r1::2, r0::2] * dVf
                else:
                    acc = acc + qf[:, r1::2, r0::2] * dVf
    return _norm(acc * dVcinv)


def _lin(n, dxf, dxc):
    gf = np.arange(n)
    ic = gf // 2
    ir = gf - ic * 2
    x = ((ir + 0.5) * dxf - dxc * 0.5) / dxc
    neg = x < 0.0
    ic = np.where(neg, ic - 1, ic)
    x = np.where(neg, x + 1.0, x)
    return ic, x


def refine(qc, dxf, normalize=True):
    """qc: (ql, [nz,] ny, nx) periodic -> (ql, [2nz,] 2ny, 2nx); without
    normalize, the plain cell-centred linear refinement of each component."""
    ndim = qc.ndim - 1
    cs = qc.shape[1:]
    fs = tuple(2 * s for s in cs)
    axes = []
    for d in range(ndim):
        ax = ndim - 1 - d   # array axis of direction d
        ic, x = _lin(fs[ax], dxf[d], 2.0 * dxf[d])
        axes.append((ic, x))
    (i0, x), (i1, y) = axes[0], axes[1]
    if ndim == 2:
        I1, I0 = np.meshgrid(i1, i0, indexing="ij")
        Y, X = np.meshgrid(y, x, indexing="ij")
        ny, nx = cs

        def a(di, dj):
            return qc[:, (I1 + dj) % ny, (I0 + di) % nx]
        q1, q2, q3, q4 = a(0, 0), a(0, 1), a(1, 0), a(1, 1)
        q = (q1 + (q3 - q1) * X) * (1.0 - Y) + (q2 + (q4 - q2) * X) * Y
    else:
        i2, z = axes[2]
        I2, I1, I0 = np.meshgrid(i2, i1, i0, indexing="ij")
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        nz, ny, nx = cs

        def a(di, dj, dk):
            return qc[:, (I2 + dk) % nz, (I1 + dj) % ny, (I0 + di) % nx]
        q1, q2, q3, q4 = a(0, 0, 0), a(0, 1, 0), a(1, 0, 0), a(1, 1, 0)
        q5, q6, q7, q8 = a(0, 0, 1), a(0, 1, 1), a(1, 0, 1), a(1, 1, 1)
        q = (((q1 + (q3 - q1) * X) * (1.0 - Y) + (q2 + (q4 - q2) * X) * Y) * (1.0 - Z) +
             ((q5 + (q7 - q5) * X) * (1.0 - Y) + (q6 + (q8 - q6) * X) * Y) * Z)
    return _norm(q) if normalize else q
