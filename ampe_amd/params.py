"""ctypes mirrors of the C ABI structs in include/ampe_hip.h and the seeded
workloads (initial conditions + model/solver parameters) the tests and the
benchmark run.

Model parameters follow the reference's input-deck keys (QuatModelParameters,
source/QuatModelParameters.cc) with the values documented per workload below.
"""
import ctypes as C

import numpy as np


class Box(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]


class CalphadPhase(C.Structure):
    _fields_ = [("gA", C.c_double), ("gB", C.c_double), ("L", C.c_double * 4)]


class CalphadGfun(C.Structure):
    _fields_ = [("nrange", C.c_int), ("Tc", C.c_double * 5), ("coef", (C.c_double * 8) * 4)]


class CalphadTdep(C.Structure):
    _fields_ = [("g", (CalphadGfun * 2) * 2), ("L", ((C.c_double * 2) * 4) * 2)]


class TrainParams(C.Structure):
    _fields_ = [("batch", C.c_longlong), ("steps", C.c_int), ("lr0", C.c_double),
                ("lr1", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double)]


class KksBatchStats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in ("iters", "max_iters", "fallbacks", "failures")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class CalphadMobility(C.Structure):
    _fields_ = [("q", (C.c_double * 2) * 2), ("qx", (C.c_double * 2) * 2),
                ("has_extra", C.c_int), ("qq", (C.c_double * 2) * 4)]


class ModelParams(C.Structure):
    _fields_ = [
        ("ndim", C.c_int),
        ("dx", C.c_double * 3),
        ("temperature", C.c_double),
        ("epsilon_phi", C.c_double),
        ("phi_well_scale", C.c_double),
        ("phi_mobility", C.c_double),
        ("phi_well_type", C.c_char),
        ("energy_interp", C.c_char),
        ("conc_interp", C.c_char),
        ("avg_type", C.c_char),
        ("flux_type", C.c_int),
        ("aniso_nu", C.c_double),
        ("aniso_knumber", C.c_int),
        ("evolve_quat", C.c_int),
        ("qlen", C.c_int),
        ("H_parameter", C.c_double),
        ("epsilon_q", C.c_double),
        ("quat_mobility", C.c_double),
        ("min_quat_mobility", C.c_double),
        ("quat_mobility_scale", C.c_double),
        ("quat_mobility_type", C.c_char),
        ("orient_interp1", C.c_char),
        ("orient_interp2", C.c_char),
        ("floor_type", C.c_char),
        ("gradient_floor", C.c_double),
        ("with_conc", C.c_int),
        ("D_liquid", C.c_double),
        ("Q0_liquid", C.c_double),
        ("D_solid", C.c_double),
        ("Q0_solid", C.c_double),
        ("conc_mobility", C.c_double),
        ("fe_model", C.c_int),
        ("vm_L", C.c_double),
        ("vm_A", C.c_double),
        ("quad_A", C.c_double * 2),
        ("quad_ceq", C.c_double * 2),
        ("quad_m", C.c_double * 2),
        ("quad_Tref", C.c_double),
        ("calphad", CalphadPhase * 2),
        ("newton_max_its", C.c_int),
        ("newton_tol", C.c_double),
        ("dilute_keq", C.c_double),
        ("dilute_m", C.c_double),
        ("dilute_Tm", C.c_double),
        ("phi_mobility_type", C.c_int),
        ("kim_interface_mobility", C.c_double),
        ("kim_antitrapping", C.c_int),
        ("conc_rhs_form", C.c_int),
        ("diff_interp", C.c_char),
        ("ebs_stencil", C.c_int),
        ("norderp", C.c_int),
        ("norderp_A", C.c_int),
        ("vm_B", C.c_double),
        ("tern_A", (C.c_double * 2) * 3),
        ("tern_ceq", (C.c_double * 2) * 3),
        ("D_solid_B", C.c_double),
        ("Q0_solid_B", C.c_double),
        ("conc_diffusion_type", C.c_int),
        ("mobility", (CalphadMobility * 2) * 2),
    ]


class ClusterParams(C.Structure):
    _fields_ = [("efficiency_tol", C.c_double), ("min_size", C.c_int * 3),
                ("max_size", C.c_int * 3), ("tag_buffer", C.c_int)]


class SolverParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("max_order", C.c_int),
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("maxl", C.c_int),
        ("max_newton", C.c_int),
        ("nlscoef", C.c_double),
        ("eplifac", C.c_double),
        ("msbp", C.c_int),
        ("lag_quat_sidegrad", C.c_int),
        ("fac_max_cycles", C.c_int),
        ("fac_sweeps", C.c_int),
        ("fac_tol", C.c_double),
        ("quat_max_cycles", C.c_int),
        ("quat_sweeps", C.c_int),
        ("precision", C.c_int),
        ("kks_guess", C.c_int),
        ("precond", C.c_int),
        ("mg_max_levels", C.c_int),
        ("mg_coarse_sweeps", C.c_int),
        ("gstype", C.c_int),
        ("graph_psolve", C.c_int),
        ("dt_init", C.c_double),
        ("adaptive", C.c_int),
        ("mg_quat_natural", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in
                ("nst", "nni", "nli", "nfe", "npe", "nps", "ncfn", "nlcf",
                 "kks_calls", "kks_cells", "kks_iters", "kks_fallbacks",
                 "kks_surrogate_starts")] + [
        ("last_acnrm", C.c_double), ("t", C.c_double), ("h", C.c_double),
        ("netf", C.c_longlong), ("q", C.c_longlong)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


FLUX_SIMPLE, FLUX_ISOTROPIC, FLUX_ANISOTROPIC = 0, 1, 2
FE_QUADRATIC, FE_CALPHAD_BINARY, FE_KKS_DILUTE, FE_QUADRATIC_TERNARY_3PH = 0, 1, 2, 3
PHI_MOB_CONSTANT, PHI_MOB_KIM = 0, 1
CONC_RHS_KKS, CONC_RHS_EBS = 0, 1
CONC_DIFF_TEMPERATURE, CONC_DIFF_COMPOSITION = 0, 1
EBS_REGULAR, EBS_ISOTROPIC, EBS_ORDER4 = 0, 1, 2
FP64, FAC_FP32 = 0, 1
GUESS_PREVIOUS, GUESS_SURROGATE = 0, 1
PRECOND_REDBLACK, PRECOND_MG = 0, 1
GS_MODIFIED, GS_CLASSICAL = 0, 1

R_GAS = 8.314472


# ---------------------------------------------------------------------------
# CALPHAD AuNi end members (thermodynamic_data/calphadAuNi.dat:7-45): piecewise
# G = a + b T + c T ln T + d2 T^2 + d3 T^3 + d7 T^7 + dm1/T + dm9 T^-9.
# ---------------------------------------------------------------------------
_AUNI = {
    ("A", "L"): dict(Tc=[298.15, 933.51, 1337.58, 1735.8, 3000.0],
                     a=[5613.147, -81023.261, 326614.987, 413.343],
                     b=[97.446385, 1012.21732, -2025.7579, 155.893158],
                     c=[-22.75455, -155.6947, 263.2523, -30.9616],
                     d2=[-0.00385924, .08756015, -0.11821685, 0.0],
                     d3=[3.79625e-7, -1.1518713e-05, 8.923845e-06, 0.0],
                     dm1=[-25097.0, 10637210.0, -67999850.0, 0.0]),
    ("A", "S"): dict(Tc=[298.15, 933.51, 1337.58, 1735.8, 3000.0],
                     a=[-6938.853, -93575.261, 314062.987, -12138.657],
                     b=[106.830495, 1021.60143, -2016.37379, 165.277268],
                     c=[-22.75455, -155.6947, 263.2523, -30.9616],
                     d2=[-0.00385924, .08756015, -0.11821685, 0.0],
                     d3=[3.79625e-7, -1.1518713e-05, 8.923845e-06, 0.0],
                     dm1=[-25097.0, 10637210.0, -67999850.0, 0.0]),
    ("B", "L"): dict(Tc=[298.15, 1728.0, 3000.0],
                     a=[11235.527, -9549.775], b=[108.457, 268.598],
                     c=[-22.096, -43.1], d2=[-.0048407, 0.0],
                     d7=[-3.82318e-21, 0.0]),
    ("B", "S"): dict(Tc=[298.15, 1728.0, 3000.0],
                     a=[-5179.159, -27840.655], b=[117.854, 279.135],
                     c=[-22.096, -43.1], d2=[-.0048407, 0.0],
                     dm9=[0.0, 1.12754e31]),
}
# Redlich-Kister L_i = L_i[0] + L_i[1] T (calphadAuNi.dat:49-63)
_AUNI_LMIX = {"L": [(9500.0, -5.429), (1614.0, 0.0)],
              "S": [(28696.0, -11.274), (-10945.0, 6.154), (2519.0, 0.0)]}


def calphad_g(species, phase, T):
    d = _AUNI[(species, phase)]
    Tc = d["Tc"]
    k = 0
    while k < len(Tc) - 2 and T >= Tc[k + 1]:
        k += 1
    g = d["a"][k] + d["b"][k] * T + d["c"][k] * T * np.log(T)
    for key, powr in (("d2", 2), ("d3", 3), ("d7", 7), ("dm1", -1), ("dm9", -9)):
        if key in d:
            g += d[key][k] * T ** powr
    return float(g)


_GFUN_KEYS = ("a", "b", "c", "d2", "d3", "d7", "dm1", "dm9")


def auni_tdep(rk=None):
    """Temperature-dependent AuNi table (CalphadTdep): end members of
    calphadAuNi.dat:7-45 per [phase L/S][species A/B] and Redlich-Kister
    L_i(T) = L_i0 + L_i1 T (calphadAuNi.dat:49-63), or the constant (L0, L1)
    of rk per phase."""
    tab = CalphadTdep()
    for ip, ph in enumerate("LS"):
        for isp, sp in enumerate("AB"):
            d = _AUNI[(sp, ph)]
            f = tab.g[ip][isp]
            f.nrange = len(d["Tc"]) - 1
            for k, tc in enumerate(d["Tc"]):
                f.Tc[k] = tc
            for k in range(f.nrange):
                for j, key in enumerate(_GFUN_KEYS):
                    f.coef[k][j] = d[key][k] if key in d else 0.0
        Ls = list(_AUNI_LMIX[ph]) if rk is None else [(v, 0.0) for v in rk[ph]]
        for i, (l0, l1) in enumerate(Ls[:4]):
            tab.L[ip][i][0] = l0
            tab.L[ip][i][1] = l1
    return tab


def seeded_rk(seed=1002):
    """Redlich-Kister (L0, L1) per phase, L0 ~ U[-2e4, 2e4], L1 ~ U[-5e3, 5e3]
    J/mol, drawn as in polycrystal_params."""
    rng = np.random.Generator(np.random.Philox(seed))
    return {ph: (float(rng.uniform(-2e4, 2e4)), float(rng.uniform(-5e3, 5e3)))
            for ph in "LS"}


def auni_phases(T, rk=None):
    """[liquid, solid] CalphadPhase at temperature T.  rk optionally overrides
    the Redlich-Kister (L0, L1) per phase: {'L': (L0, L1), 'S': (L0, L1)}."""
    out = (CalphadPhase * 2)()
    for i, ph in enumerate("LS"):
        out[i].gA = calphad_g("A", ph, T)
        out[i].gB = calphad_g("B", ph, T)
        Ls = [a + b * T for a, b in _AUNI_LMIX[ph]]
        if rk is not None:
            Ls = list(rk[ph])
        Ls = (Ls + [0.0] * 4)[:4]
        for j in range(4):
            out[i].L[j] = Ls[j]
    return out


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

def solver_defaults(**kw):
    s = SolverParams(dt=1.0e-3, max_order=2, rtol=1.0e-6, atol=1.0e-4, maxl=5,
                     max_newton=3, nlscoef=0.1, eplifac=0.05, msbp=1,
                     lag_quat_sidegrad=1, fac_max_cycles=1, fac_sweeps=2,
                     fac_tol=-1.0, quat_max_cycles=1, quat_sweeps=2,
                     precision=FP64, kks_guess=GUESS_PREVIOUS,
                     precond=PRECOND_REDBLACK, mg_max_levels=0, mg_coarse_sweeps=16,
                     gstype=GS_MODIFIED, graph_psolve=0, adaptive=0)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def two_grain_params(ndim=2, **kw):
    """Quadratic-KKS two-grain model: values of tests/TwoGrainsQuadratic/
    2d.input:23-58 (H, eps_q, eps_phi, mobilities, well scale, hbsm
    interpolation, quadratic wells, D/Q), lengths rescaled by 1/0.05 so that
    dx = 1 (eps -> eps/L, H -> H/L, D -> D/L^2 with L = 0.05 um)."""
    L = 0.05
    p = ModelParams()
    p.ndim = ndim
    for d in range(3):
        p.dx[d] = 1.0
    p.temperature = 873.0
    p.epsilon_phi = 0.165 / L
    p.phi_well_scale = 0.4125
    p.phi_mobility = 200.0
    p.phi_well_type = b"d"
    p.energy_interp = b"h"
    p.conc_interp = b"h"
    p.avg_type = b"a"
    p.flux_type = FLUX_SIMPLE
    p.aniso_nu = 0.0
    p.aniso_knumber = 4
    p.evolve_quat = 1
    p.qlen = 4
    p.H_parameter = 0.001 / L
    p.epsilon_q = 0.1 / L
    p.quat_mobility = 200.0
    p.min_quat_mobility = 1.0e-6
    p.quat_mobility_scale = 1.0
    p.quat_mobility_type = b"p"
    p.orient_interp1 = b"q"
    p.orient_interp2 = b"c"
    p.floor_type = b"m"
    p.gradient_floor = 1.0e-2
    p.with_conc = 1
    p.D_liquid = 5.6e4 / L ** 2
    p.Q0_liquid = 55329.0
    p.D_solid = 1.3e8 / L ** 2
    p.Q0_solid = 156377.0
    p.conc_mobility = 1.0
    p.fe_model = FE_QUADRATIC
    p.vm_L = 1.5e-5
    p.vm_A = 1.5e-5
    p.quad_A[0], p.quad_A[1] = 1.0e4, 1.0e4
    p.quad_ceq[0], p.quad_ceq[1] = 0.05, 0.10
    p.quad_m[0], p.quad_m[1] = 0.0, 0.0
    p.quad_Tref = 873.0
    p.calphad = auni_phases(873.0)
    p.newton_max_its = 50
    p.newton_tol = 1.0e-12
    p.diff_interp = b"l"
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def polycrystal_params(ndim=3, seed=1002, **kw):
    """Binary CALPHAD polycrystal at T = 1500 K in the reference's physical
    units (um, pJ, s): interface energy sigma = 0.372678 J/m^2 and width
    delta = 0.035 um (tests/KKScomposition/2d.input:32-35) converted as
    QuatModelParameters does (eps_phi = sqrt(6 sigma delta), w = 3 sigma /
    (16 delta), source/QuatModelParameters.cc:842-845), phase mobility 6.4 and
    AuNi diffusivities D_S = 0.125, D_L = 1224.23 um^2/s (same deck), molar
    volume 7.68e-6 m^3/mol, pbg interpolation, AuNi end members
    (thermodynamic_data/calphadAuNi.dat) with seeded Redlich-Kister
    L0 ~ U[-2e4, 2e4], L1 ~ U[-5e3, 5e3] J/mol per phase.  Orientation
    coefficients of tests/TwoGrainsQuadratic/2d.input.  dx = 0.02 um.
    No penalty term (the reference's penalty solver is not compiled)."""
    rk = seeded_rk(seed)
    sigma, delta = 0.372678, 0.035
    p = two_grain_params(ndim=ndim)
    for d in range(3):
        p.dx[d] = 0.02
    p.temperature = 1500.0
    p.epsilon_phi = float(np.sqrt(6.0 * sigma * delta))
    p.phi_well_scale = (3.0 * sigma / delta) / 16.0
    p.phi_mobility = 6.4
    p.energy_interp = b"p"
    p.conc_interp = b"p"
    p.H_parameter = 0.001
    p.epsilon_q = 0.1
    p.quat_mobility = 200.0
    p.fe_model = FE_CALPHAD_BINARY
    p.vm_L = 7.68e-6
    p.vm_A = 7.68e-6
    p.calphad = auni_phases(1500.0, rk)
    p.D_liquid = 1224.23
    p.Q0_liquid = 0.0
    p.D_solid = 0.125
    p.Q0_solid = 0.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def dilute_params(ndim=2, **kw):
    """Dilute binary solution (KKSdiluteBinary) with the CuNi values of
    tests/KKSdiluteBinaryKKS/test.input and tests/DiluteCuNi/2d.input:
    k_e = 0.7965, liquidus slope -310.9 K, T_m = 1728 K, T = 1710 K, linear
    concentration interpolation, on the polycrystal's interface and
    orientation parameters."""
    p = polycrystal_params(ndim=ndim)
    p.fe_model = FE_KKS_DILUTE
    p.temperature = 1710.0
    p.dilute_keq = 0.7965
    p.dilute_m = -310.9
    p.dilute_Tm = 1728.0
    p.conc_interp = b"l"
    p.vm_L = p.vm_A = 6.8812e-6
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def kim_interface_mobility(vm_liquid, Tm, liquidus_slope, keq, beta):
    """Interface mobility [um^3/(pJ s)] from the kinetics coefficient beta [s/m]
    (kinetics_coefficient), source/QuatModelParameters.cc:1156-1166."""
    mu = vm_liquid / (R_GAS * Tm)
    mu *= 1.0e6
    mu *= -1.0 * liquidus_slope
    mu /= 1.0 - keq
    mu /= beta
    return mu


def random_unit_quaternions(rng, n):
    """Uniform on S3 (Shoemake's method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    s1, s2 = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([s1 * np.sin(2 * np.pi * u2), s1 * np.cos(2 * np.pi * u2),
                     s2 * np.sin(2 * np.pi * u3), s2 * np.cos(2 * np.pi * u3)])


def two_grain_state(n=128, seed=1001, radius=32.0, width=4.0, ndim=2):
    """Disk of radius 32 cells (phi = 1 inside, tanh interface of width 4) in
    a phi = 0 matrix; c = 0.45 in the solid, 0.40 in the liquid, plus uniform
    noise +-0.005; one S3-uniform quaternion per grain.  Arrays are (comp,)
    (nz,) ny, nx with i fastest."""
    rng = np.random.Generator(np.random.Philox(seed))
    shape = (n, n) if ndim == 2 else (n, n, n)
    idx = np.indices(shape).astype(np.float64) + 0.5
    r = np.sqrt(sum((x - n / 2.0) ** 2 for x in idx))
    phi = 0.5 * (1.0 - np.tanh((r - radius) / (width / 2.0)))
    c = np.where(phi >= 0.5, 0.45, 0.40) + rng.uniform(-0.005, 0.005, shape)
    qs = random_unit_quaternions(rng, 2)
    inside = (phi >= 0.5)
    q = np.empty((4,) + shape)
    for m in range(4):
        q[m] = np.where(inside, qs[m, 0], qs[m, 1])
    return phi, c, q


def polycrystal_state(n=256, seed=1002, ngrains=64, solid_fraction=0.7,
                      width=3.0, ndim=3):
    """Periodic Voronoi polycrystal: ngrains seeds U[0,n)^d, each with an
    S3-uniform quaternion.  The smooth field d(x) = distance to the nearest
    seed selects the liquid pockets: cells above the solid_fraction-quantile
    of d are liquid (phi = 0), with tanh interfaces of the given width.
    c ~ U[0.3, 0.5] per cell."""
    rng = np.random.Generator(np.random.Philox(seed))
    shape = (n,) * ndim
    seeds = rng.uniform(0, n, (ngrains, ndim))
    quats = random_unit_quaternions(rng, ngrains)
    idx = [np.arange(n, dtype=np.float64) + 0.5 for _ in range(ndim)]
    grids = np.meshgrid(*idx[::-1], indexing="ij")[::-1]  # x fastest
    best = np.full(shape, np.inf)
    label = np.zeros(shape, dtype=np.int32)
    for g in range(ngrains):
        d2 = np.zeros(shape)
        for d in range(ndim):
            t = np.abs(grids[d] - seeds[g, d])
            t = np.minimum(t, n - t)
            d2 += t * t
        closer = d2 < best
        best = np.where(closer, d2, best)
        label = np.where(closer, g, label)
    dist = np.sqrt(best)
    thr = np.quantile(dist, solid_fraction)
    phi = 0.5 * (1.0 - np.tanh((dist - thr) / (width / 2.0)))
    c = rng.uniform(0.3, 0.5, shape)
    q = quats[:, label]
    return phi, c, q


def kks_composition_deck():
    """The reference's KKScomposition regression (tests/KKScomposition/
    2d.input, tests/KKScomposition/test2d.py): 2-D 64 x 64 cells on
    [0, 1.8]^2 um with zero-slope boundaries (periodic_dimension = 0, 0;
    "slope" 0 conditions), AuNi CALPHAD KKS (calphadAuNi.dat, Redlich-Kister
    terms of the database), sigma = 0.372678 J/m^2, delta = 0.035 um,
    M_phi = 6.4, pbg interpolation, arithmetic face averages, D_S = 0.125,
    D_L = 1224.23 um^2/s, molar volume 7.68e-6 m^3/mol, KKS Newton max_its 50,
    no orientation field; temperature 1450 K cooled at -200 K/s towards
    1220 K; end_time 0.3 s, atol 1e-4 (rtol = atol 1e-2), PhaseSysSolver
    red-black with 5 sweeps and no early exit.  Initial state of test2d.py:
    make_nuclei.py --nx 64 --ny 64 --nz 1 -r 16 --center0 "0, 0, 0" -c 0.25
    --concentration-in 0.096.  The test passes when the integral of c stays
    within 1e-4 of its initial value and the solid volume fraction at the
    end time is 0.32 +- 0.01.

    Returns (model params, ramp dict, initial state, checks dict)."""
    from . import initial
    sigma, delta = 0.372678, 0.035
    n = 64
    p = two_grain_params(ndim=2)
    for d in range(3):
        p.dx[d] = 1.8 / n
    p.temperature = 1450.0
    p.epsilon_phi = float(np.sqrt(6.0 * sigma * delta))
    p.phi_well_scale = (3.0 * sigma / delta) / 16.0
    p.phi_mobility = 6.4
    p.energy_interp = b"p"
    p.conc_interp = b"p"
    p.avg_type = b"a"
    p.evolve_quat = 0
    p.fe_model = FE_CALPHAD_BINARY
    p.vm_L = p.vm_A = 7.68e-6
    p.calphad = auni_phases(1450.0)
    p.D_liquid, p.Q0_liquid = 1224.23, 0.0
    p.D_solid, p.Q0_solid = 0.125, 0.0
    p.newton_max_its = 50
    ramp = dict(T0=1450.0, dTdt=-200.0, T_target=1220.0, tab=auni_tdep())
    ic = initial.make_nuclei(nx=n, ny=n, nz=1, radius=16, center0=(0.0, 0.0, 0.0),
                             nomconc=[0.25], conc_in=[0.096])
    state = (ic["phase"][0], ic["conc"][0][0], None)
    checks = dict(end_time=0.3, conc_integral_tol=1.0e-4, solid_fraction=0.32,
                  solid_fraction_tol=1.0e-2, max_cycles=900)
    return p, ramp, state, checks


# CALPHAD atomic mobilities of Au (species 0) and Ni (species 1) in the
# liquid and in fcc (thermodynamic_data/calphadAuNi.dat:66-106,
# MobilityParameters): (value, prefactor) of qA, qB, q0AB, q1AB; the absent
# q2AB, q3AB are (0, 1).
_AUNI_MOB = {
    ("L", 0): dict(qA=(-35921.87, 3.1514e-8), qB=(-46353.23, 5.9576e-8),
                   q0AB=(0.0, 1.0), q1AB=(0.0, 1.0)),
    ("A", 0): dict(qA=(-176600.0, 1.002878e-05), qB=(-306500.0, 4.49132e-2),
                   q0AB=(232300.0, 7.6881435e-08), q1AB=(15900.0, 1.0)),
    ("L", 1): dict(qA=(-35921.87, 3.6471e-8), qB=(-46353.23, 6.8948e-8),
                   q0AB=(0.0, 1.0), q1AB=(0.0, 1.0)),
    ("A", 1): dict(qA=(-188800.0, 1.702455873e-05), qB=(-287000.0, 2.25994e-4),
                   q0AB=(136200.0, 1.0), q1AB=(-42800.0, 1.0)),
}


def set_mobility(m, qA, qB, q0AB=(0.0, 1.0), q1AB=(0.0, 1.0), q2AB=(0.0, 1.0),
                 q3AB=(0.0, 1.0), extraA=None, extraB=None):
    """Fill one CalphadMobility record (CALPHADMobility::initialize,
    source/CALPHADMobility.cc:12-153: absent Redlich-Kister terms (0, 1);
    four-value qA/qB entries carry the T ln T and 1/T coefficients)."""
    for x, v in enumerate((qA, qB)):
        m.q[x][0], m.q[x][1] = v
    m.has_extra = 1 if (extraA is not None or extraB is not None) else 0
    for x, v in enumerate((extraA, extraB)):
        m.qx[x][0], m.qx[x][1] = v if v is not None else (0.0, 0.0)
    for n, v in enumerate((q0AB, q1AB, q2AB, q3AB)):
        m.qq[n][0], m.qq[n][1] = v


def auni_mobility(p):
    """p.mobility[phase][species] from calphadAuNi.dat's MobilityParameters
    (PhaseL -> 0, PhaseA -> 1; Species0 = Au, Species1 = Ni)."""
    for ip, ph in enumerate("LA"):
        for sp in range(2):
            set_mobility(p.mobility[ip][sp], **_AUNI_MOB[(ph, sp)])
    return p


def single_grain_growth_auni_deck():
    """The reference's SingleGrainGrowthAuNi regression (tests/
    SingleGrainGrowthAuNi/2d.input, test2d.py): the KKScomposition set-up
    (2-D 64 x 64 cells on [0, 1.8]^2 um, zero-slope boundaries, AuNi CALPHAD
    KKS, M_phi = 6.4, pbg interpolation, arithmetic averages, molar volume
    7.68e-6, KKS Newton max_its 50, 1450 K cooled at -200 K/s towards
    1220 K, make_nuclei.py -r 16 --center0 "0, 0, 0" -c 0.25
    --concentration-in 0.096) with epsilon_phi = 0.25 (pJ/um)^1/2,
    phi_well_scale = 2.5 pJ/um^3, the EBS composition form and its default
    composition-dependent diffusion (the database's atomic mobilities,
    linear diffusion weights, source/QuatModelParameters.cc:238-250, 1014-1016),
    atol 1e-5.  test2d.py: the integral of c within 1e-4 of its initial value,
    solid volume fraction 0.32 +- 0.01 after t = 0.3 s.

    Returns (model params, ramp dict, initial state, checks dict)."""
    p, ramp, state, checks = kks_composition_deck()
    p.epsilon_phi = 0.25
    p.phi_well_scale = 2.5
    p.conc_rhs_form = CONC_RHS_EBS
    p.ebs_stencil = EBS_REGULAR
    p.diff_interp = b"l"
    p.conc_diffusion_type = CONC_DIFF_COMPOSITION
    auni_mobility(p)
    return p, ramp, state, checks


def kks_composition_solver(dt=1.0e-3, **kw):
    """Solver settings of the KKScomposition deck on the fixed-step BDF
    integrator (the reference's CVODE chooses its own steps up to end_time)."""
    base = dict(dt=dt, dt_init=1.0e-6, atol=1.0e-4, rtol=1.0e-6, maxl=10, max_newton=5, fac_sweeps=5,
                fac_max_cycles=1, fac_tol=-1.0)
    base.update(kw)
    return solver_defaults(**base)


def one_grain_quadratic_deck(ndim=2):
    """The reference's OneGrainQuadratic regression (tests/OneGrainQuadratic/
    2d.input, test2d.py): 2-D 64 x 64 periodic cells on [0, 3.2]^2 um,
    quadratic KKS wells (A = 1e4, c_eq = 0.05 / 0.10, m = 0, T_ref = 1000 K),
    EBS composition flux with temperature-dependent diffusion (D_S = 1.3e8,
    Q_S = 156377; D_L = 5.6e4, Q_L = 55329; linear phase weights),
    sigma = 0.1 J/m^2, delta = 0.045 um, M_phi = 200, hbsm interpolation,
    arithmetic averages, molar volumes 1.5e-5 m^3/mol, no orientation field;
    temperature 873 K cooled at -20 K/s towards 573 K; end_time 0.25 s,
    atol 1e-4.  Initial state of test2d.py: make_nuclei.py --nx 64 --ny 64
    --nz 1 -r 8 --concentration-in 0.1 --concentration-out 0.06 --ngrains 1.
    The test passes when the solid volume fraction at the end time is
    0.21 +- 0.01 (test2d.py:39-46).  ndim = 3: 48^3 cells on [0, 2.4]^3 um,
    make_nuclei.py --nx 48 --ny 48 --nz 48 -r 8, end_time 0.15 s, solid
    fraction 0.14 +- 0.01 (3d.input, test3d.py:36-46).

    Returns (model params, ramp dict, initial state, checks dict)."""
    from . import initial
    sigma, delta = 0.1, 0.045
    n = 64 if ndim == 2 else 48
    p = two_grain_params(ndim=ndim)
    for d in range(3):
        p.dx[d] = 0.05
    p.temperature = 873.0
    p.epsilon_phi = float(np.sqrt(6.0 * sigma * delta))
    p.phi_well_scale = (3.0 * sigma / delta) / 16.0
    p.phi_mobility = 200.0
    p.energy_interp = b"h"
    p.conc_interp = b"h"
    p.avg_type = b"a"
    p.evolve_quat = 0
    p.fe_model = FE_QUADRATIC
    p.vm_L = p.vm_A = 1.5e-5
    p.quad_A[0], p.quad_A[1] = 1.0e4, 1.0e4
    p.quad_ceq[0], p.quad_ceq[1] = 0.05, 0.10
    p.quad_m[0], p.quad_m[1] = 0.0, 0.0
    p.quad_Tref = 1000.0
    p.D_liquid, p.Q0_liquid = 5.6e4, 55329.0
    p.D_solid, p.Q0_solid = 1.3e8, 156377.0
    p.conc_rhs_form = CONC_RHS_EBS
    p.diff_interp = b"l"
    ramp = dict(T0=873.0, dTdt=-20.0, T_target=573.0, tab=None)
    ic = initial.make_nuclei(nx=n, ny=n, nz=1 if ndim == 2 else n, radius=8, conc_in=[0.1],
                             conc_out=[0.06])
    if ndim == 2:
        state = (ic["phase"][0], ic["conc"][0][0], None)
        checks = dict(end_time=0.25, solid_fraction=0.21, solid_fraction_tol=1.0e-2)
    else:
        state = (ic["phase"], ic["conc"][0], None)
        checks = dict(end_time=0.15, solid_fraction=0.14, solid_fraction_tol=1.0e-2)
    checks["shape"] = state[0].shape
    return p, ramp, state, checks


def two_grains_quadratic_deck(ndim=2):
    """The reference's TwoGrainsQuadratic regression (tests/TwoGrainsQuadratic/
    2d.input, test2d.py): the OneGrainQuadratic setting with an orientation
    field (H = 0.001, epsilon_orient = 0.1, orient_mobility = 200, the
    reference's defaults for the rest), epsilon_phi = 0.165,
    phi_well_scale = 0.4125, the KKS composition flux, T_ref = 873 K; two
    nuclei of radii 8 and 12 cells with random orientations (make_nuclei.py
    --nx 64 --ny 64 --nz 1 -r 8 --concentration-in 0.1
    --concentration-out 0.06 --ngrains 2 -q 4).  The test passes when the
    solid volume fraction at t = 0.25 s is 0.27 +- 0.01 and the grain volumes
    reported every 0.05 s (cells with phi >= 0.85, GrainDiagnostics) range
    from 0.3025 to 1.9175 um^2, each +- 0.01 (test2d.py:30-75).  ndim = 3:
    64 x 64 x 48 cells on [0, 3.2] x [0, 3.2] x [0, 2.4] um, end_time 0.08 s,
    diagnostics every 0.01 s, solid fraction 0.13 +- 0.01, grain volumes
    up to 2.0 +- 0.01 and down to 0.179 +- 0.001 um^3 (3d.input,
    test3d.py:28-70).

    Returns (model params, ramp dict, initial state, checks dict)."""
    from . import initial
    n = 64
    nz = 1 if ndim == 2 else 48
    p = two_grain_params(ndim=ndim)
    for d in range(3):
        p.dx[d] = 0.05
    p.temperature = 873.0
    p.epsilon_phi = 0.165
    p.phi_well_scale = 0.4125
    p.phi_mobility = 200.0
    p.H_parameter = 0.001
    p.epsilon_q = 0.1
    p.quat_mobility = 200.0
    p.gradient_floor = 1.0e-2
    p.D_liquid, p.Q0_liquid = 5.6e4, 55329.0
    p.D_solid, p.Q0_solid = 1.3e8, 156377.0
    p.quad_Tref = 873.0
    ramp = dict(T0=873.0, dTdt=-20.0, T_target=573.0, tab=None)
    ic = initial.make_nuclei(nx=n, ny=n, nz=nz, radius=8, ngrains=2, conc_in=[0.1],
                             conc_out=[0.06], qlen=4)
    if ndim == 2:
        state = (ic["phase"][0], ic["conc"][0][0], ic["quat"][:, 0])
        checks = dict(end_time=0.25, solid_fraction=0.27, solid_fraction_tol=1.0e-2,
                      grain_threshold=0.85, grain_interval=0.05, grain_vmax=1.9175,
                      grain_vmin=0.3025, grain_tol=1.0e-2, grain_tol_min=1.0e-2)
    else:
        state = (ic["phase"], ic["conc"][0], ic["quat"])
        checks = dict(end_time=0.08, solid_fraction=0.13, solid_fraction_tol=1.0e-2,
                      grain_threshold=0.85, grain_interval=0.01, grain_vmax=2.0,
                      grain_vmin=0.179, grain_tol=1.0e-2, grain_tol_min=1.0e-3)
    checks["shape"] = state[0].shape
    return p, ramp, state, checks


def sintering_deck():
    """The reference's Sintering regression (tests/Sintering/2d.input,
    test2d.py): 2-D 64 x 64 cells on [0, 3.2]^2 um, norderp = 5 order
    parameters -- two grains of solid A, two of solid B, the liquid (pores)
    last -- with the ternary three-phase quadratic model (A = 1e4 J/mol for
    every phase and species, c_eq L = (0, 0), A = (1, 0), B = (0, 1)), EBS
    composition flux with temperature-dependent diffusion (D_A = D_B = 1e4,
    D_L = 0, Q = 0, linear weights), sigma = 0.1 J/m^2, delta = 0.045 um,
    M_phi = 200, arithmetic averages, molar volumes 1.5e-5 m^3/mol,
    T = 1000 K, periodic; atol 1e-4 (rtol atol/100), end_time 0.006 s in at
    most 150 steps.  Initial state: make_multi_spheres.py --nx 64 --ny 64
    --nz 1 --concentration-A 1.,0. --concentration-B 0.,1.
    --concentration-out 0.,0. --spheres 4spheres.csv (spheres of radius 12
    at (20, 20), (44, 20) [A] and (20, 44), (44, 44) [B]); the liquid order
    parameter is 1 - the sum of the others (FieldsInitializer,
    source/FieldsInitializer.cc:441-456).  The test passes when, at the end
    time, the largest and smallest volume fractions over the five order
    parameters are 0.586 and 0.118 (+- 0.001, test2d.py:59-69).

    Returns (model params, initial state (phi[5], c[2]), checks dict)."""
    from . import initial
    sigma, delta = 0.1, 0.045
    n = 64
    p = two_grain_params(ndim=2)
    for d in range(3):
        p.dx[d] = 0.05
    p.temperature = 1000.0
    p.epsilon_phi = float(np.sqrt(6.0 * sigma * delta))
    p.phi_well_scale = (3.0 * sigma / delta) / 16.0
    p.phi_mobility = 200.0
    p.flux_type = FLUX_SIMPLE
    p.energy_interp = b"p"
    p.conc_interp = b"l"
    p.avg_type = b"a"
    p.evolve_quat = 0
    p.with_conc = 1
    p.fe_model = FE_QUADRATIC_TERNARY_3PH
    p.norderp, p.norderp_A = 5, 2
    p.vm_L = p.vm_A = p.vm_B = 1.5e-5
    for a in range(3):
        for k in range(2):
            p.tern_A[a][k] = 1.0e4
    ceq = ((0.0, 0.0), (1.0, 0.0), (0.0, 1.0))
    for a in range(3):
        for k in range(2):
            p.tern_ceq[a][k] = ceq[a][k]
    p.D_liquid, p.Q0_liquid = 0.0, 0.0
    p.D_solid, p.Q0_solid = 1.0e4, 0.0
    p.D_solid_B, p.Q0_solid_B = 1.0e4, 0.0
    p.conc_rhs_form = CONC_RHS_EBS
    p.diff_interp = b"l"
    p.ebs_stencil = EBS_REGULAR
    spheres = [(20, 20, 16, 12, "A"), (44, 20, 16, 12, "A"), (20, 44, 16, 12, "B"),
               (44, 44, 16, 12, "B")]
    ic = initial.make_multi_spheres(n, n, spheres, conc_a=(1.0, 0.0), conc_b=(0.0, 1.0),
                                    conc_out=(0.0, 0.0))
    grains = ic["phase"][:, 0].astype(np.float64)
    phi = np.concatenate([grains, (1.0 - grains.sum(axis=0))[None]])
    c = ic["conc"][:, 0].astype(np.float64)
    checks = dict(end_time=0.006, max_steps=150, vmax=0.586, vmin=0.118, tol=1.0e-3)
    return p, (phi, c), checks

