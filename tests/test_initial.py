"""The restated initial-condition generator (utils/make_nuclei.py) against
closed forms: lattice-point counts of the nucleus, the exterior composition
that meets the nominal one, single-precision storage, the tanh profile and
the two-nucleus layout."""
import numpy as np

from ampe_amd import initial


def test_kks_composition_nucleus():
    """tests/KKScomposition/test2d.py: 64 x 64, r = 16 at (0, 0, 0), -c 0.25,
    --concentration-in 0.096 (cell centres at z = 0.5 add 0.25 to d^2)."""
    ic = initial.make_nuclei(nx=64, ny=64, nz=1, radius=16, center0=(0.0, 0.0, 0.0),
                             nomconc=[0.25], conc_in=[0.096])
    phi = ic["phase"][0]
    # lattice points (i + 1/2)^2 + (j + 1/2)^2 + 1/4 < 256, counted row by row
    count = sum(int(np.sum((np.arange(64) + 0.5) ** 2 < 256 - 0.25 - (j + 0.5) ** 2))
                for j in range(64))
    assert phi.sum() == count == ic["vs"]
    assert set(np.unique(phi)) <= {0.0, 1.0}
    vol = 64 * 64
    co = (0.25 * vol - 0.096 * count) / (vol - count)
    assert ic["conc_out"][0] == float(np.float32(co))
    c = ic["conc"][0][0]
    assert np.all(c[phi == 1.0] == float(np.float32(0.096)))
    assert np.all(c[phi == 0.0] == float(np.float32(co)))
    # the nominal composition is met up to single-precision rounding
    assert abs(c.mean() - 0.25) < 1e-7


def test_tanh_profile_and_double_precision():
    ic = initial.make_nuclei(nx=32, ny=32, nz=1, radius=8, width=1.5, double_precision=True)
    phi = ic["phase"][0]
    j, i = 20, 22
    d = np.sqrt((i + 0.5 - 16) ** 2 + (j + 0.5 - 16) ** 2 + (0.5 - 1) ** 2) - 8
    assert abs(phi[j, i] - 0.5 * (1 + np.tanh(-d / 3.0))) < 1e-15
    # cells beyond 8 widths keep the exterior phase
    assert phi[0, 0] == 0.0
    assert abs(ic["vs"] - phi.sum()) < 1e-9


def test_two_nuclei():
    ic = initial.make_nuclei(nx=60, ny=60, nz=1, radius=5, ngrains=2)
    phi = ic["phase"][0]
    # centres (20, 20) and (40, 40), radii 5 and 7.5, z offset 0.5 - 0.5 = 0
    assert phi[19, 19] == 1.0 and phi[39, 39] == 1.0
    assert phi[39, 39 + 7] == 1.0 and phi[19, 19 + 6] == 0.0


def test_random_orientations_two_nuclei():
    """make_nuclei.py -q 4 --ngrains 2: Shoemake quaternions on the generator's
    discrete grid (u1 in steps of 0.1, u2 and u3 in steps of 0.05), unit
    norm, and every cell carries one of the two."""
    ic = initial.make_nuclei(nx=64, ny=64, nz=1, radius=8, ngrains=2, conc_in=[0.1],
                             conc_out=[0.06], qlen=4)
    qs = np.array(ic["grain_quats"])
    assert np.allclose((qs ** 2).sum(axis=1), 1.0)
    for qv in qs:
        u1 = qv[2] ** 2 + qv[3] ** 2
        assert abs(u1 * 10 - round(u1 * 10)) < 1e-9
    q = ic["quat"][:, 0]
    g0 = np.all(q == qs[0].astype(np.float32).astype(np.float64)[:, None, None], axis=0)
    g1 = np.all(q == qs[1].astype(np.float32).astype(np.float64)[:, None, None], axis=0)
    assert np.all(g0 | g1) and g0[20, 20] and g1[43, 43]
