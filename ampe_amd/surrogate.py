"""Training of the KKS surrogate on the GPU.

The training set is (c, phi, T) drawn uniformly over the given ranges and the
(cL, cS) the FP64 batch KKS Newton (ampe_kks_batch, default start) converges
to; the network (ampe_surrogate: 3 -> 64 -> 64 -> 2, tanh, BF16 matrix-core
kernels with FP32 accumulation) is trained by the library's own forward,
backward and Adam kernels.  Device memory comes from torch; all arithmetic
runs in libampe_hip.so.
"""
import time

import numpy as np

from ampe_amd import hip as H

C_RANGE = (0.05, 0.95)
PHI_RANGE = (0.0, 1.0)
T_RANGE = (1200.0, 1600.0)


def scaling(c_range=C_RANGE, phi_range=PHI_RANGE, T_range=T_RANGE):
    """Inputs mapped to [-1, 1]; outputs (cL, cS) from [0, 1] to [-1, 1]."""
    lo = np.array([c_range[0], phi_range[0], T_range[0]])
    hi = np.array([c_range[1], phi_range[1], T_range[1]])
    return (0.5 * (lo + hi), 2.0 / (hi - lo), (0.5, 0.5), (0.5, 0.5))


def draw_cells(n, seed, c_range=C_RANGE, phi_range=PHI_RANGE, T_range=T_RANGE):
    """(c, phi, T) uniform over the ranges (Philox stream of the seed)."""
    rng = np.random.Generator(np.random.Philox(seed))
    return (rng.uniform(*c_range, n), rng.uniform(*phi_range, n), rng.uniform(*T_range, n))


def kks_targets(p, tab, x, dev):
    """FP64 Newton solutions (cL, cS) for device inputs x[3][n]."""
    import torch
    n = x.shape[1]
    cl = torch.zeros(n, dtype=torch.float64, device=dev)
    cs = torch.zeros_like(cl)
    st = H.kks_batch(p, tab, x[0], x[1], x[2], cl, cs)
    if st["failures"]:
        raise H.AmpeError(f"KKS Newton failed on {st['failures']} training cells")
    return torch.stack([cl, cs]).contiguous(), st


def train_kks_surrogate(p, tab, dev, n_samples=10_000_000, seed=1003, steps=20000,
                        batch=8192, lr0=3e-3, lr1=3e-5, net_seed=5):
    """Returns (surrogate, info).  p: binary CALPHAD model parameters
    (interpolation, Newton tolerance); tab: temperature-dependent table."""
    import torch
    t0 = time.perf_counter()
    x = torch.from_numpy(np.stack(draw_cells(n_samples, seed))).to(dev)
    tgt, st = kks_targets(p, tab, x, dev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    s = H.Surrogate(3, 64, 2, seed=net_seed)
    s.set_scaling(*scaling())
    hist = s.train(x, tgt, batch=batch, steps=steps, lr0=lr0, lr1=lr1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    info = dict(samples=n_samples, steps=steps, batch=batch, data_s=t1 - t0, train_s=t2 - t1,
                loss_first=float(np.mean(hist[:20])), loss_last=float(np.mean(hist[-100:])),
                data_newton_iters_per_cell=st["iters"] / n_samples)
    return s, info


def surrogate_error(s, p, tab, dev, n=1 << 16, seed=1004):
    """Max and RMS |guess - Newton solution| over fresh cells (concentration
    units), from the forward pass on normalised inputs."""
    import torch
    x = torch.from_numpy(np.stack(draw_cells(n, seed))).to(dev)
    tgt, _ = kks_targets(p, tab, x, dev)
    sh, sc, osh, osc = s.scaling
    xn = ((x.float() - torch.tensor(sh, dtype=torch.float32, device=dev)[:, None])
          * torch.tensor(sc, dtype=torch.float32, device=dev)[:, None]).contiguous()
    y = torch.zeros((2, n), dtype=torch.float32, device=dev)
    s.forward(xn, y)
    guess = y.double() * torch.tensor(osc, dtype=torch.float64, device=dev)[:, None] \
        + torch.tensor(osh, dtype=torch.float64, device=dev)[:, None]
    err = (guess - tgt).abs()
    return dict(max_abs=float(err.max()), rms=float(err.pow(2).mean().sqrt()))
