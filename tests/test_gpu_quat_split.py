"""The FP32 multigrid stores the orientation block's vectors row
parity-split (inside each ghosted row the even columns, then the odd ones)
unless mg_quat_natural = 1.  The layout changes where values sit, not which
operations run on them, so the integrator must give bit-identical states and
counters either way: one periodic patch with three multigrid levels (Krylov
kernels forming the FP32 input and reading its solution, captured solve),
two uneven periodic patches (modified Gram-Schmidt: the flat conversion
kernels; two orientation cycles: the split residual feeds the next cycle),
and a bounded 2-D level with odd row lengths (one level, zero-slope
ghosts)."""
import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu


def _box(lo, hi):
    b = P.Box()
    b.lo[:] = list(lo) + [0] * (3 - len(lo))
    b.hi[:] = list(hi) + [0] * (3 - len(hi))
    return b


def _compact(a, boxes, ndim):
    """Interior values of the global array a (([nz,] ny, nx), or with a
    leading component axis) in the level's compact order, patch by patch."""
    parts = []
    for b in boxes:
        if ndim == 3:
            sl = (slice(b.lo[2], b.hi[2] + 1), slice(b.lo[1], b.hi[1] + 1),
                  slice(b.lo[0], b.hi[0] + 1))
        else:
            sl = (slice(b.lo[1], b.hi[1] + 1), slice(b.lo[0], b.hi[0] + 1))
        if a.ndim > ndim:
            parts.append(a[(slice(None),) + sl].reshape(a.shape[0], -1))
        else:
            parts.append(a[sl].reshape(-1))
    return np.concatenate(parts, axis=-1)


CASES = {
    "3d-periodic-3levels-classical-graph": dict(
        ndim=3, n=(32, 32, 32), boxes=[((0, 0, 0), (31, 31, 31))], periodic=True,
        solver=dict(gstype=P.GS_CLASSICAL, graph_psolve=1, mg_max_levels=3)),
    "3d-two-patches-modified-2cycles": dict(
        ndim=3, n=(32, 24, 16), boxes=[((0, 0, 0), (11, 23, 15)), ((12, 0, 0), (31, 23, 15))],
        periodic=True,
        solver=dict(gstype=P.GS_MODIFIED, quat_max_cycles=2, fac_tol=-1.0, mg_max_levels=3)),
    "2d-bounded-odd-rows": dict(
        ndim=2, n=(45, 38), boxes=[((0, 0), (44, 37))], periodic=False,
        solver=dict(gstype=P.GS_CLASSICAL, mg_max_levels=3)),
}


def _run(cfg, natural, steps):
    from ampe_amd import hip as H
    ndim = cfg["ndim"]
    nx, ny = cfg["n"][0], cfg["n"][1]
    nz = cfg["n"][2] if ndim == 3 else 1
    big = max(cfg["n"])
    p = P.polycrystal_params(ndim=ndim, seed=1002)
    assert p.evolve_quat == 1 and p.qlen == 4  # the split applies to the orientation block
    phi, c, q = P.polycrystal_state(n=big, seed=1002, ngrains=6, ndim=ndim)
    if ndim == 3:
        phi, c, q = phi[:nz, :ny, :nx], c[:nz, :ny, :nx], q[:, :nz, :ny, :nx]
    else:
        phi, c, q = phi[:ny, :nx], c[:ny, :nx], q[:, :ny, :nx]
    boxes = [_box(lo, hi) for lo, hi in cfg["boxes"]]
    s = P.solver_defaults(dt=5e-8, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8,
                          precision=P.FAC_FP32, precond=P.PRECOND_MG, fac_sweeps=1,
                          quat_sweeps=1, mg_coarse_sweeps=4, mg_quat_natural=natural,
                          **cfg["solver"])
    dom = _box((0, 0, 0), (nx - 1, ny - 1, nz - 1)) if cfg["periodic"] else None
    sim = H.Sim(p, s, boxes, periodic=dom)
    sim.set_state(_compact(phi, boxes, ndim), _compact(c, boxes, ndim), _compact(q, boxes, ndim))
    out = []
    for _ in range(steps):
        rc, st = sim.advance(1)
        assert rc == 0, st
        out.append((dict(st), [a.copy() for a in sim.get_state((sim.ncell,))]))
    return out


@pytest.mark.parametrize("case", sorted(CASES))
def test_split_layout_matches_natural(dev, case):
    cfg = CASES[case]
    print("case", case, flush=True)
    split = _run(cfg, 0, 3)
    natural = _run(cfg, 1, 3)
    for step, ((ss, fs), (sn, fn)) in enumerate(zip(split, natural)):
        for k in ("nst", "nni", "nli", "nfe", "npe", "nps", "ncfn", "nlcf"):
            assert ss[k] == sn[k], (case, step, k, ss[k], sn[k])
        assert ss["t"] == sn["t"]
        for name, a, b in zip(("phi", "c", "q", "cl", "cs"), fs, fn):
            assert np.array_equal(a, b), (case, step, name, float(np.max(np.abs(a - b))))
    print(case, "nli", split[-1][0]["nli"], "nps", split[-1][0]["nps"])


def _halved(boxes):
    out = []
    for b in boxes:
        c = P.Box()
        for a in range(3):
            c.lo[a] = b.lo[a] // 2
            c.hi[a] = b.lo[a] // 2 + (b.hi[a] - b.lo[a] + 1) // 2 - 1
        out.append(c)
    return out


def _domain(lev, buf, shape):
    """Interior values of a ghosted level array as the domain array."""
    out = np.zeros(shape)
    for p, b in enumerate(lev.boxes):
        blk = lev.ghost_view(buf, p)
        if len(shape) == 3:
            out[b.lo[2]:b.hi[2] + 1, b.lo[1]:b.hi[1] + 1, b.lo[0]:b.hi[0] + 1] = blk[1:-1, 1:-1, 1:-1]
        else:
            out[b.lo[1]:b.hi[1] + 1, b.lo[0]:b.hi[0] + 1] = blk[1:-1, 1:-1]
    return out


@pytest.mark.parametrize("shape,split", [((48, 64), None), ((48, 64), 20),
                                         ((16, 24, 32), None), ((16, 24, 32), 12)])
def test_red_residual_restriction(dev, shape, split):
    """After red-black sweeps ending with the black colour the residual of the
    orientation block vanishes on black cells (to rounding), so restricting
    the red residuals alone (red_only = 1, the multigrid's path on periodic
    levels) gives the coarse right-hand side of the full residual."""
    import ctypes as C
    import torch
    from ampe_amd import hip as H
    ndim = len(shape)
    ext = list(shape[::-1]) + [1] * (3 - ndim)
    cuts = [(0, ext[0] - 1)] if split is None else [(0, split - 1), (split, ext[0] - 1)]
    boxes = []
    for lo, hi in cuts:
        b = _box((lo, 0, 0), (hi, ext[1] - 1, ext[2] - 1))
        boxes.append(b)
    full = H.boxes_of(shape, ndim)
    fine = H.Level(ndim, boxes, 1, periodic=full)
    cb = _halved([full])[0]
    coarse = H.Level(ndim, _halved(boxes), 1, periodic=cb)
    rng = np.random.default_rng(7)
    fc = torch.from_numpy(-rng.uniform(0.5, 1.5, fine.nside_tot)).to(dev)
    sm = torch.from_numpy(fine.to_ghosted(rng.uniform(0.5, 1.5, shape))).to(dev)
    ncomp = 4
    rhs = torch.from_numpy(np.concatenate([fine.to_ghosted(rng.normal(size=shape))
                                           for _ in range(ncomp)])).to(dev)
    z = torch.from_numpy(np.concatenate([fine.to_ghosted(rng.normal(size=shape))
                                         for _ in range(ncomp)])).to(dev)
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    H.check(H.lib().ampe_ghost_fill(fine.h, ptr(sm), 1, None))
    dx = (C.c_double * 3)(0.5, 0.25, 0.75)
    H.check(H.lib().ampe_quat_smooth(fine.h, ptr(fc), ptr(sm), 0.3, dx, ncomp, ptr(rhs), ptr(z),
                                     1, 0, None))
    res = torch.zeros_like(z)
    b_full = torch.zeros(ncomp * coarse.ncg, dtype=torch.float64, device=dev)
    b_red = torch.zeros_like(b_full)
    H.check(H.lib().ampe_quat_residual_restrict(coarse.h, fine.h, ptr(fc), ptr(sm), 0.3, dx,
                                                ncomp, ptr(rhs), ptr(z), ptr(res), ptr(b_full),
                                                0, None))
    H.check(H.lib().ampe_quat_residual_restrict(coarse.h, fine.h, ptr(fc), ptr(sm), 0.3, dx,
                                                ncomp, ptr(rhs), ptr(z), None, ptr(b_red), 1,
                                                None))
    # black residuals are rounding noise next to the red ones
    r = res.cpu().numpy().reshape(ncomp, -1)
    inter = np.stack([_domain(fine, r[m], shape) for m in range(ncomp)])
    idx = np.indices(shape).sum(axis=0)
    black = (idx % 2) == 1
    assert np.abs(inter[:, black]).max() < 1e-12 * np.abs(inter[:, ~black]).max()
    bf, br = b_full.cpu().numpy(), b_red.cpu().numpy()
    assert np.abs(br - bf).max() < 1e-12 * np.abs(bf).max()
    # a bounded level is refused
    fine_b = H.Level(ndim, boxes, 1, periodic=None)
    coarse_b = H.Level(ndim, _halved(boxes), 1, periodic=None)
    assert H.lib().ampe_quat_residual_restrict(coarse_b.h, fine_b.h, ptr(fc), ptr(sm), 0.3, dx,
                                               ncomp, ptr(rhs), ptr(z), None, ptr(b_red), 1,
                                               None) != 0


@pytest.mark.parametrize("shape,split", [((48, 64), 20), ((16, 24, 32), None), ((16, 24, 32), 12)])
@pytest.mark.parametrize("variant", ["phase", "conc"])
def test_red_residual_restriction_elliptic(dev, shape, split, variant):
    """The same for the phase block (C and M cell fields, constant D) and the
    composition block (D face field): after red-black sweeps the black
    residuals are rounding noise, and the red-only coarse right-hand side
    (ampe_fac_residual_restrict, red_only = 1) matches the full one."""
    import ctypes as C
    import torch
    from ampe_amd import hip as H
    ndim = len(shape)
    ext = list(shape[::-1]) + [1] * (3 - ndim)
    cuts = [(0, ext[0] - 1)] if split is None else [(0, split - 1), (split, ext[0] - 1)]
    boxes = [_box((lo, 0, 0), (hi, ext[1] - 1, ext[2] - 1)) for lo, hi in cuts]
    full = H.boxes_of(shape, ndim)
    fine = H.Level(ndim, boxes, 1, periodic=full)
    coarse = H.Level(ndim, _halved(boxes), 1, periodic=_halved([full])[0])
    rng = np.random.default_rng(9)
    op = H.EllipticOp()
    keep = []
    if variant == "phase":
        Cf = torch.from_numpy(fine.to_ghosted(rng.uniform(1.0, 2.0, shape))).to(dev)
        Mf = torch.from_numpy(fine.to_ghosted(rng.uniform(0.5, 1.5, shape))).to(dev)
        op.C, op.M = Cf.data_ptr(), Mf.data_ptr()
        op.D_const, op.C_const, op.M_const = -0.7, 1.0, 1.0
        keep += [Cf, Mf]
    else:
        Df = torch.from_numpy(-rng.uniform(0.2, 2.0, fine.nside_tot)).to(dev)
        op.D = Df.data_ptr()
        op.D_const, op.C_const, op.M_const = 0.0, 1.0, 1.3
        keep.append(Df)
    op.dx = (C.c_double * 3)(0.5, 0.25, 0.75)
    rhs = torch.from_numpy(fine.to_ghosted(rng.normal(size=shape))).to(dev)
    u = torch.from_numpy(fine.to_ghosted(rng.normal(size=shape))).to(dev)
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    H.check(H.lib().ampe_fac_smooth(fine.h, C.byref(op), ptr(rhs), ptr(u), 1, -1.0, None, None))
    res = torch.zeros_like(u)
    b_full = torch.zeros(coarse.ncg, dtype=torch.float64, device=dev)
    b_red = torch.zeros_like(b_full)
    H.check(H.lib().ampe_fac_residual_restrict(coarse.h, fine.h, C.byref(op), ptr(rhs), ptr(u),
                                               ptr(res), ptr(b_full), 0, None))
    H.check(H.lib().ampe_fac_residual_restrict(coarse.h, fine.h, C.byref(op), ptr(rhs), ptr(u),
                                               None, ptr(b_red), 1, None))
    inter = _domain(fine, res.cpu().numpy(), shape)
    black = (np.indices(shape).sum(axis=0) % 2) == 1
    assert np.abs(inter[black]).max() < 1e-12 * np.abs(inter[~black]).max()
    bf, br = b_full.cpu().numpy(), b_red.cpu().numpy()
    assert np.abs(br - bf).max() < 1e-12 * np.abs(bf).max()
