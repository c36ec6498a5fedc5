"""The FP32 preconditioner solve captured into a HIP graph and replayed
(graph_psolve = 1) gives the integrator the same states and counters as the
eager launches, step after step.  Every step replays the graph once per
Krylov iteration, with the integrator's own eager kernels and host
synchronisations between the replays; the first steps change the BDF order
and so gamma, which forces new captures."""
import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu


def _sim(p, s, state, shape):
    from ampe_amd import hip as H
    box = H.boxes_of(shape, len(shape))
    sim = H.Sim(p, s, [box], periodic=box)
    sim.set_state(*state)
    return sim


CASES = {
    "3d-calphad-mg": dict(shape=(32, 32, 32), mg=True),
    "3d-calphad-redblack": dict(shape=(32, 32, 32), mg=False),
    "2d-quadratic-mg": dict(shape=(64, 64), mg=True),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_graph_psolve_matches_eager(dev, case):
    cfg = CASES[case]
    shape = cfg["shape"]
    if len(shape) == 3:
        p = P.polycrystal_params(seed=1002)
        state = P.polycrystal_state(n=shape[0], seed=1002, ngrains=8)
        base = dict(dt=5e-8, atol=1e-6, rtol=1e-8)
    else:
        p = P.two_grain_params(ndim=2, quad_ceq=(0.40, 0.45))
        state = P.two_grain_state(n=shape[0])
        base = dict(dt=1e-3, atol=1e-6, rtol=1e-8)
    base.update(maxl=10, max_newton=5, precision=P.FAC_FP32, gstype=P.GS_CLASSICAL,
                fac_sweeps=1, quat_sweeps=1)
    if cfg["mg"]:
        base.update(precond=P.PRECOND_MG, mg_max_levels=3, mg_coarse_sweeps=2)
    eager = _sim(p, P.solver_defaults(**base), state, shape)
    graph = _sim(p, P.solver_defaults(graph_psolve=1, **base), state, shape)
    for step in range(4):
        re, se = eager.advance(1)
        rg, sg = graph.advance(1)
        assert re == 0 and rg == 0
        for k in ("nst", "nni", "nli", "nfe", "npe", "nps", "ncfn", "nlcf"):
            assert se[k] == sg[k], (step, k, se[k], sg[k])
        assert se["t"] == sg["t"]
        fe = eager.get_state(shape)
        fg = graph.get_state(shape)
        for name, a, b in zip(("phi", "c", "q", "cl", "cs"), fe, fg):
            assert np.array_equal(a, b), (case, step, name, float(np.max(np.abs(a - b))))
    print(case, "steps", se["nst"], "solves", se["nps"])
