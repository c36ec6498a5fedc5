"""GPU checks of the KKS surrogate path through the C ABI: BF16 matrix-core
forward and loss gradient against the oracle (same BF16 operand rounding,
FP64 accumulation), the per-cell-temperature KKS batch against the oracle,
training on FP64 Newton solutions, what the trained surrogate saves over the
default start, and the integrator with the surrogate guess."""
import numpy as np
import pytest

from ampe_amd import params as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NIN, HID, NOUT = 3, 64, 2
IN_SHIFT, IN_SCALE = (0.5, 0.5, 1400.0), (1 / 0.45, 2.0, 1 / 200.0)
OUT_SHIFT, OUT_SCALE = (0.5, 0.5), (0.5, 0.5)


def _cells(n, seed):
    rng = np.random.Generator(np.random.Philox(seed))
    return rng.uniform(0.05, 0.95, n), rng.uniform(0.0, 1.0, n), rng.uniform(1200.0, 1600.0, n)


def _blocks(g):
    o = [0, HID * NIN, HID, HID * HID, HID, NOUT * HID, NOUT]
    edges = np.cumsum(o)
    return [g[edges[i]:edges[i + 1]] for i in range(6)]


@pytest.mark.parametrize("n", [64, 4000])
def test_forward_matches_oracle(dev, n):
    from ampe_amd import hip as H
    import torch
    s = H.Surrogate(NIN, HID, NOUT, seed=21)
    w = s.get_params()
    rng = np.random.default_rng(1)
    x = rng.uniform(-1.2, 1.2, (NIN, n)).astype(np.float32)
    xd = torch.from_numpy(x).to(dev)
    yd = torch.zeros((NOUT, n), dtype=torch.float32, device=dev)
    s.forward(xd, yd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().astype(np.float64)
    ye = O.mlp_forward(w, x, NIN, HID, NOUT, emulate=1)
    y0 = O.mlp_forward(w, x, NIN, HID, NOUT, emulate=0)
    scale = np.max(np.abs(ye))
    print("forward max|gpu-emul|", np.max(np.abs(y - ye)), "max|gpu-fp64|", np.max(np.abs(y - y0)))
    assert np.max(np.abs(y - ye)) < 2e-3 * scale
    assert np.linalg.norm(y - ye) < 1e-3 * np.linalg.norm(ye)
    assert np.linalg.norm(y - y0) < 3e-2 * np.linalg.norm(y0)


@pytest.mark.parametrize("n", [16, 3001])
def test_loss_grad_matches_oracle(dev, n):
    from ampe_amd import hip as H
    import torch
    s = H.Surrogate(NIN, HID, NOUT, seed=22)
    w = s.get_params()
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (NIN, n)).astype(np.float32)
    t = rng.uniform(-1, 1, (NOUT, n)).astype(np.float32)
    g = torch.zeros(s.nparams, dtype=torch.float32, device=dev)
    loss = s.loss_grad(torch.from_numpy(x).to(dev), torch.from_numpy(t).to(dev), g)
    torch.cuda.synchronize()
    gg = g.cpu().numpy().astype(np.float64)
    le, ge = O.mlp_loss_grad(w, x, t, NIN, HID, NOUT, emulate=1)
    l0, g0 = O.mlp_loss_grad(w, x, t, NIN, HID, NOUT, emulate=0)
    print("loss", loss, le, l0, "grad rel |gpu-emul|", np.linalg.norm(gg - ge) / np.linalg.norm(ge))
    assert abs(loss - le) < 1e-4 * le
    assert np.linalg.norm(gg - ge) < 5e-3 * np.linalg.norm(ge)
    for a, b in zip(_blocks(gg), _blocks(ge)):
        assert np.linalg.norm(a - b) < 1e-2 * np.linalg.norm(b) + 1e-9
    assert np.linalg.norm(gg - g0) < 5e-2 * np.linalg.norm(g0)


def test_kks_batch_default_start_matches_oracle(dev):
    from ampe_amd import hip as H
    import torch
    p = P.polycrystal_params(seed=1002)
    tab = P.auni_tdep(P.seeded_rk())
    n = 20011
    c, phi, T = _cells(n, 31)
    cl_o, cs_o, it_o, nf_o = O.kks_batch(p, tab, c, phi, T)
    td = [torch.from_numpy(a).to(dev) for a in (c, phi, T)]
    cl = torch.zeros(n, dtype=torch.float64, device=dev)
    cs = torch.zeros_like(cl)
    it = torch.zeros(n, dtype=torch.int32, device=dev)
    st = H.kks_batch(p, tab, *td, cl, cs, it)
    cl, cs, it = cl.cpu().numpy(), cs.cpu().numpy(), it.cpu().numpy()
    print("kks batch", st, "oracle mean its", it_o.mean())
    assert nf_o == 0 and st["failures"] == 0 and st["fallbacks"] == 0
    assert st["iters"] == int(it.sum()) and st["max_iters"] == int(it.max())
    assert np.mean(it == it_o) > 0.999
    assert np.max(np.abs(cl - cl_o)) < 1e-12 and np.max(np.abs(cs - cs_o)) < 1e-12


@pytest.fixture(scope="module")
def trained(dev):
    """Surrogate trained on 2^18 FP64-Newton solutions (GPU batch KKS from the
    default start) of the config-3 input distribution."""
    from ampe_amd import hip as H
    import torch
    p = P.polycrystal_params(seed=1002)
    tab = P.auni_tdep(P.seeded_rk())
    n = 1 << 18
    c, phi, T = _cells(n, 41)
    x = torch.from_numpy(np.stack([c, phi, T])).to(dev)
    cl = torch.zeros(n, dtype=torch.float64, device=dev)
    cs = torch.zeros_like(cl)
    st = H.kks_batch(p, tab, x[0], x[1], x[2], cl, cs)
    assert st["failures"] == 0
    tgt = torch.stack([cl, cs]).contiguous()
    s = H.Surrogate(NIN, HID, NOUT, seed=5)
    s.set_scaling(IN_SHIFT, IN_SCALE, OUT_SHIFT, OUT_SCALE)
    hist = s.train(x, tgt, batch=4096, steps=3000, lr0=3e-3, lr1=1e-4)
    return p, tab, s, hist


def test_training_reduces_loss(trained):
    _, _, _, hist = trained
    first, last = float(np.mean(hist[:20])), float(np.mean(hist[-50:]))
    print("training loss first", first, "last", last)
    assert np.all(np.isfinite(hist))
    assert last < first / 100.0


def test_surrogate_guess_saves_newton_iterations(trained, dev):
    """Fresh cells: the surrogate start converges every cell to the same root
    with fewer Newton updates than the (0.5, 0.5) start."""
    from ampe_amd import hip as H
    import torch
    p, tab, s, _ = trained
    n = 1 << 16
    c, phi, T = _cells(n, 43)
    td = [torch.from_numpy(a).to(dev) for a in (c, phi, T)]
    out = {}
    for name, guess in (("default", None), ("surrogate", s)):
        cl = torch.zeros(n, dtype=torch.float64, device=dev)
        cs = torch.zeros_like(cl)
        it = torch.zeros(n, dtype=torch.int32, device=dev)
        st = H.kks_batch(p, tab, *td, cl, cs, it, guess=guess, probe_its=3, probe_tol=1e-8)
        out[name] = (cl.cpu().numpy(), cs.cpu().numpy(), it.cpu().numpy(), st)
    d, g = out["default"], out["surrogate"]
    print("default", d[3], "surrogate", g[3])
    assert d[3]["failures"] == 0 and g[3]["failures"] == 0
    assert g[3]["iters"] < d[3]["iters"]
    same = (np.abs(g[0] - d[0]) < 1e-10) & (np.abs(g[1] - d[1]) < 1e-10)
    assert np.mean(same) > 0.999
    # every surrogate-started cell is a KKS solution to the Newton tolerance
    cl_o, cs_o, it_o, nf = O.kks_batch(p, tab, c, phi, T, gl=g[0], gs=g[1])
    assert nf == 0 and it_o.max() == 0


def test_integrator_with_surrogate_guess(trained):
    """A few BDF steps of the 3-D CALPHAD polycrystal with the surrogate KKS
    guess (every cell, and only where the previous-step guess is poor) agree
    with the previous-step guess run: every KKS solve converges to 1e-12."""
    from ampe_amd import hip as H
    _, _, s, _ = trained
    n = 16
    p = P.polycrystal_params(seed=1002)
    phi, c, q = P.polycrystal_state(n=n, seed=1002, ngrains=4)
    box = H.boxes_of((n, n, n), 3)
    res = {}
    for name, guess, switch in (("previous", P.GUESS_PREVIOUS, None),
                                ("surrogate_all", P.GUESS_SURROGATE, 0.0),
                                ("surrogate_switch", P.GUESS_SURROGATE, 1e-2)):
        sv = P.solver_defaults(dt=2e-7, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8,
                               kks_guess=guess)
        sim = H.Sim(p, sv, [box], periodic=box)
        if switch is not None:
            sim.set_surrogate(s, probe_its=3, probe_tol=1e-8, switch_tol=switch)
        sim.set_state(phi, c, q)
        rc, st = sim.advance(3)
        assert rc == 0
        res[name] = (sim.get_state((n, n, n)), st)
    a, sa = res["previous"]
    ncg = (n + 2) ** 3
    for name in ("surrogate_all", "surrogate_switch"):
        b, sb = res[name]
        print(name, "kks iters", sb["kks_iters"], "vs previous", sa["kks_iters"],
              "surrogate starts", sb["kks_surrogate_starts"], "fallbacks", sb["kks_fallbacks"])
        for x, y in zip(a[:3], b[:3]):
            assert np.linalg.norm(x - y) < 1e-9 * np.linalg.norm(x)
    assert res["surrogate_all"][1]["kks_surrogate_starts"] == \
        res["surrogate_all"][1]["kks_calls"] * ncg
    assert sa["kks_surrogate_starts"] == 0
