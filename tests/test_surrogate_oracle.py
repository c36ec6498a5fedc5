"""CPU checks of the oracle for the KKS surrogate path: BF16 rounding, the
temperature-dependent CALPHAD table, the per-cell-temperature KKS batch and
the MLP forward / backward (no GPU)."""
import numpy as np
import pytest
import torch

from ampe_amd import params as P
from oracle import oracle as O

R = P.R_GAS
NIN, HID, NOUT = 3, 64, 2


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0,
                        rng.uniform(-1e-3, 1e-3, 2000).astype(np.float32)])
    # exact ties: mantissa bit pattern ...8000 (round to even)
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties, np.zeros(1, np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = np.array([O.lib().o_bf16(float(v)) for v in x], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("T", [300.0, 933.51, 1200.0, 1337.58, 1450.0, 1600.0, 1728.0, 1800.0,
                               2500.0])
def test_calphad_table_matches_database(T):
    """G(T) of the C table vs the database evaluation in ampe_amd.params (itself
    checked against tests/golden/calphadAuNi.dat)."""
    tab = P.auni_tdep()
    ph = O.calphad_at(tab, T)
    ref = P.auni_phases(T)
    for i in range(2):
        for a, b in ((ph[i].gA, ref[i].gA), (ph[i].gB, ref[i].gB)):
            assert abs(a - b) <= 1e-12 * max(1.0, abs(b))
        for j in range(4):
            assert abs(ph[i].L[j] - ref[i].L[j]) <= 1e-9


def _cells(n, seed):
    rng = np.random.Generator(np.random.Philox(seed))
    return rng.uniform(0.05, 0.95, n), rng.uniform(0.0, 1.0, n), rng.uniform(1200.0, 1600.0, n)


def test_kks_batch_invariants_and_constant_T_path():
    rk = P.seeded_rk()
    p = P.polycrystal_params(seed=1002)
    tab = P.auni_tdep(rk)
    c, phi, T = _cells(4000, 11)
    cl, cs, it, nf = O.kks_batch(p, tab, c, phi, T)
    assert nf == 0 and it.min() >= 0
    h = phi ** 3 * (10 - 15 * phi + 6 * phi ** 2)
    assert np.max(np.abs((1 - h) * cl + h * cs - c)) < 1e-12
    for k in range(0, 4000, 97):
        ph = O.calphad_at(tab, T[k])
        p.temperature = T[k]
        for i in range(2):
            p.calphad[i] = ph[i]
        dmu = O.lib().o_dfenergy(p, cl[k], 0) - O.lib().o_dfenergy(p, cs[k], 1)
        assert abs(dmu) / (R * T[k]) < 1e-12
        # same cell through the constant-temperature path (o_kks_cell)
        xl, xs = np.array([0.5]), np.array([0.5])
        r = O.lib().o_kks_cell(p, c[k], phi[k], O.ptr(xl), O.ptr(xs))
        assert r == it[k] and xl[0] == cl[k] and xs[0] == cs[k]


def test_kks_batch_probe_restart():
    """A guess far from the root fails the probe and restarts from (0.5, 0.5):
    same root, iterations = probe + default-start count."""
    p = P.polycrystal_params(seed=1002)
    tab = P.auni_tdep(P.seeded_rk())
    c, phi, T = _cells(500, 12)
    cl0, cs0, it0, _ = O.kks_batch(p, tab, c, phi, T)
    bad = np.full(500, 0.999)
    cl1, cs1, it1, nf = O.kks_batch(p, tab, c, phi, T, gl=bad, gs=bad, probe_its=0,
                                    probe_tol=0.0)
    assert nf == 0
    assert np.array_equal(cl1, cl0) and np.array_equal(cs1, cs0)
    assert np.array_equal(it1, it0)
    # exact guess: converges with zero updates
    cl2, cs2, it2, _ = O.kks_batch(p, tab, c, phi, T, gl=cl0, gs=cs0)
    assert it2.max() == 0


def _net(seed=3):
    rng = np.random.default_rng(seed)
    npar = HID * NIN + HID + HID * HID + HID + NOUT * HID + NOUT
    w = (rng.standard_normal(npar) * 0.3).astype(np.float32)
    x = rng.uniform(-1, 1, (NIN, 40)).astype(np.float32)
    t = rng.uniform(-1, 1, (NOUT, 40)).astype(np.float32)
    return w, x, t


def _torch_loss(w, x, t):
    wt = torch.tensor(w.astype(np.float64), requires_grad=True)
    o = 0
    W1 = wt[o:o + HID * NIN].view(HID, NIN); o += HID * NIN
    b1 = wt[o:o + HID]; o += HID
    W2 = wt[o:o + HID * HID].view(HID, HID); o += HID * HID
    b2 = wt[o:o + HID]; o += HID
    W3 = wt[o:o + NOUT * HID].view(NOUT, HID); o += NOUT * HID
    b3 = wt[o:o + NOUT]
    X = torch.tensor(x.astype(np.float64)).T
    h1 = torch.tanh(X @ W1.T + b1)
    h2 = torch.tanh(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    loss = ((y - torch.tensor(t.astype(np.float64)).T) ** 2).sum() / X.shape[0]
    loss.backward()
    return loss.item(), wt.grad.numpy(), y.detach().numpy().T


def test_mlp_fp64_matches_torch_autograd():
    w, x, t = _net()
    loss, g = O.mlp_loss_grad(w, x, t, NIN, HID, NOUT, emulate=0)
    tl, tg, ty = _torch_loss(w, x, t)
    y = O.mlp_forward(w, x, NIN, HID, NOUT, emulate=0)
    assert np.max(np.abs(y - ty)) < 1e-12
    assert abs(loss - tl) < 1e-12 * tl
    assert np.max(np.abs(g - tg)) < 1e-12 * np.max(np.abs(tg))


def test_mlp_fp64_gradient_finite_differences():
    w, x, t = _net(5)
    _, g = O.mlp_loss_grad(w, x, t, NIN, HID, NOUT, emulate=0)
    rng = np.random.default_rng(9)
    idx = np.unique(np.concatenate([rng.integers(0, w.size, 60), [0, HID * NIN, w.size - 1]]))
    for j in idx:
        eps = np.float32(1e-3)
        wp, wm = w.copy(), w.copy()
        wp[j] += eps
        wm[j] -= eps
        lp, _ = O.mlp_loss_grad(wp, x, t, NIN, HID, NOUT, emulate=0)
        lm, _ = O.mlp_loss_grad(wm, x, t, NIN, HID, NOUT, emulate=0)
        fd = (lp - lm) / (float(wp[j]) - float(wm[j]))
        assert abs(fd - g[j]) <= 1e-4 * max(1e-3, abs(g[j])), (j, fd, g[j])


def test_mlp_bf16_emulation_close_to_fp64():
    w, x, t = _net(6)
    l0, g0 = O.mlp_loss_grad(w, x, t, NIN, HID, NOUT, emulate=0)
    l1, g1 = O.mlp_loss_grad(w, x, t, NIN, HID, NOUT, emulate=1)
    assert abs(l1 - l0) < 2e-2 * l0
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)
    assert not np.array_equal(g1, g0)
