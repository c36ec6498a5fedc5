"""The NetCDF initial-data reader (ampe_amd/fields_io.py) against the
semantics of the reference's FieldsInitializer (source/FieldsInitializer.cc,
NETCDF4 branch) on classic files written here: variable names and their
fallbacks, the 'phase'-for-every-order and 1-minus-sum rules, dimension and
qlen checks, the precision of the first phase variable, uniform fills, and a
generator-made deck read back and stepped by the oracle integrator exactly as
when the generator's fields are passed directly."""
import os

import numpy as np
import pytest

from ampe_amd import fields_io as F
from ampe_amd import initial as I
from ampe_amd import params as P


def _raw(path, dims, variables):
    """A classic file with the given (name, code, array) variables."""
    from scipy.io import netcdf_file
    f = netcdf_file(path, "w", version=2)
    for name, n in dims.items():
        f.createDimension(name, n)
    for name, code, a in variables:
        v = f.createVariable(name, code, ("z", "y", "x"))
        v[:] = a
    f.close()


def test_round_trip_double_and_float(tmp_path):
    rng = np.random.default_rng(1)
    ph = rng.uniform(0, 1, (6, 5, 4))
    q = rng.normal(size=(4, 6, 5, 4))
    c = rng.uniform(0, 1, (1, 6, 5, 4))
    T = rng.uniform(900, 1000, (6, 5, 4))
    p = str(tmp_path / "a.nc")
    F.write_fields(p, ph, conc=c, quat=q, temperature=T, double_precision=True)
    out = F.read_fields(p, (4, 5, 6), read_t=True)
    assert np.array_equal(out["phase"][0], ph) and np.array_equal(out["quat"], q)
    assert np.array_equal(out["conc"], c) and np.array_equal(out["temperature"], T)
    F.write_fields(p, ph, conc=c, quat=q)
    out = F.read_fields(p, (4, 5, 6))
    assert np.array_equal(out["phase"][0], ph.astype(np.float32).astype(np.float64))
    assert not np.array_equal(out["phase"][0], ph)
    assert "temperature" not in out


def test_names_and_fallbacks(tmp_path):
    p = str(tmp_path / "b.nc")
    dims = dict(x=3, y=2, z=2)
    a, b, c = (np.full((2, 2, 3), v) for v in (0.25, 0.5, 0.125))
    # one species: 'concentration', else 'concentration0'
    _raw(p, dims, [("phase", "d", a), ("concentration", "d", b)])
    assert F.read_fields(p, (3, 2, 2), qlen=0)["conc"][0, 0, 0, 0] == 0.5
    _raw(p, dims, [("phase", "d", a), ("concentration0", "d", c)])
    assert F.read_fields(p, (3, 2, 2), qlen=0)["conc"][0, 0, 0, 0] == 0.125
    # two species: 'concentration0', 'concentration1' (no bare name)
    _raw(p, dims, [("phase", "d", a), ("concentration0", "d", b), ("concentration1", "d", c)])
    out = F.read_fields(p, (3, 2, 2), qlen=0, ncompositions=2)
    assert out["conc"][0, 0, 0, 0] == 0.5 and out["conc"][1, 0, 0, 0] == 0.125
    _raw(p, dims, [("phase", "d", a), ("concentration", "d", b)])
    with pytest.raises(F.FieldsFileError, match="concentration0"):
        F.read_fields(p, (3, 2, 2), qlen=0, ncompositions=2)
    # 'phase' serves every order when present; else phase<i>
    _raw(p, dims, [("phase", "d", a), ("phase1", "d", b), ("concentration", "d", c)])
    out = F.read_fields(p, (3, 2, 2), qlen=0, nphases=2)
    assert np.all(out["phase"][0] == 0.25) and np.all(out["phase"][1] == 0.25)
    _raw(p, dims, [("phase0", "d", a), ("phase1", "d", b), ("concentration", "d", c)])
    out = F.read_fields(p, (3, 2, 2), qlen=0, nphases=2, norder=3)
    assert np.all(out["phase"][1] == 0.5)
    # one order more than the file: 1 minus the sum of the others
    assert np.all(out["phase"][2] == 1.0 - (0.25 + 0.5))
    _raw(p, dims, [("phase1", "d", b), ("concentration", "d", c)])
    with pytest.raises(F.FieldsFileError, match="phase0"):
        F.read_fields(p, (3, 2, 2), qlen=0, nphases=2)


def test_dimension_and_qlen_checks(tmp_path):
    p = str(tmp_path / "c.nc")
    ph = np.zeros((2, 2, 3))
    F.write_fields(p, ph, conc=[ph], quat=[ph, ph], double_precision=True)
    with pytest.raises(F.FieldsFileError, match="nx_file=3"):
        F.read_fields(p, (4, 2, 2), qlen=2)
    with pytest.raises(F.FieldsFileError, match="qlen_file=2, QLEN=4"):
        F.read_fields(p, (3, 2, 2), qlen=4)
    out = F.read_fields(p, (3, 2, 2), qlen=2)
    assert out["quat"].shape == (2, 2, 2, 3)
    # 2-D: the file must have one plane (nz_prob = 1)
    with pytest.raises(F.FieldsFileError, match="nz_file=2"):
        F.read_fields(p, (3, 2), qlen=2)
    # uniform fills when the file is not read
    out = F.read_fields(p, (3, 2, 2), qlen=4, read_q=False, uniform_q=[1, 0, 0, 0],
                        read_c=False, uniform_c=[0.3], uniform_t=900.0)
    assert np.all(out["quat"][0] == 1) and np.all(out["conc"] == 0.3)
    assert np.all(out["temperature"] == 900.0)
    with pytest.raises(F.FieldsFileError, match="need 4 values"):
        F.read_fields(p, (3, 2, 2), qlen=4, read_q=False, uniform_q=[1, 0])


def test_precision_follows_the_first_phase_variable(tmp_path):
    """A float phase makes the reader buffer every field in single precision
    (FieldsInitializer.cc:320-334), so a double composition is rounded."""
    p = str(tmp_path / "d.nc")
    dims = dict(x=3, y=2, z=1)
    c = np.full((1, 2, 3), 0.1)
    _raw(p, dims, [("phase", "f", np.zeros((1, 2, 3), np.float32)), ("concentration", "d", c)])
    out = F.read_fields(p, (3, 2), qlen=0)
    assert out["conc"][0, 0, 0] == float(np.float32(0.1)) != 0.1
    _raw(p, dims, [("phase", "d", np.zeros((1, 2, 3))), ("concentration", "d", c)])
    assert F.read_fields(p, (3, 2), qlen=0)["conc"][0, 0, 0] == 0.1


def test_netcdf4_files_are_refused(tmp_path):
    p = tmp_path / "e.nc"
    p.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(F.FieldsFileError, match="nccopy -k classic"):
        F.read_fields(str(p), (3, 2))
    p.write_bytes(b"not a file")
    with pytest.raises(F.FieldsFileError, match="not a NetCDF"):
        F.read_fields(str(p), (3, 2))


def test_generator_deck_read_back_steps_like_the_generator(tmp_path):
    """The OneGrainQuadratic deck's initial fields (make_nuclei.py, single
    precision) written as make_nuclei.py lays them out, read back and stepped
    by the oracle integrator: the same trajectory as from the generator's
    arrays."""
    from oracle import oracle as O
    p, ramp, st, ck = P.one_grain_quadratic_deck()
    ic = I.make_nuclei(nx=64, ny=64, nz=1, radius=8, conc_in=[0.1], conc_out=[0.06])
    path = str(tmp_path / "nuclei.nc")
    F.write_fields(path, ic["phase"][0], conc=[cs[0] for cs in ic["conc"]])
    fields = F.read_fields(path, (64, 64), qlen=0)
    assert np.array_equal(fields["phase"][0], st[0]) and np.array_equal(fields["conc"][0], st[1])
    s = P.kks_composition_solver(dt=2e-3, atol=1e-6, rtol=1e-8)
    out = []
    for phi, c in ((st[0], st[1]), (fields["phase"][0], fields["conc"][0])):
        sim = O.Sim(p, s, (64, 64, 1))
        sim.set_temperature_ramp(ramp["T0"], ramp["dTdt"], ramp["T_target"], ramp["tab"])
        sim.set_state(phi, c, None)
        rc, stt = sim.advance(3)
        assert rc == 0
        out.append(sim.get_state()[:2])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_patch_values_follow_the_level_order():
    b0, b1 = P.Box(), P.Box()
    b0.lo[:], b0.hi[:] = [0, 0, 0], [2, 3, 0]
    b1.lo[:], b1.hi[:] = [3, 0, 0], [6, 3, 0]
    a = np.arange(4 * 7, dtype=float).reshape(4, 7)
    v = F.patch_values(a, [b0, b1], 2)
    assert np.array_equal(v[:12], a[:, :3].ravel()) and np.array_equal(v[12:], a[:, 3:].ravel())
    v2 = F.patch_values(np.stack([a, -a]), [b0, b1], 2)
    assert v2.shape == (2, 28) and np.array_equal(v2[1], -v)
