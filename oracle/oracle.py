"""ctypes binding of the FP64 C oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ampe_amd.params import CalphadGfun, CalphadPhase, CalphadTdep, ModelParams, SolverParams, Stats
from ampe_amd.params import FE_QUADRATIC_TERNARY_3PH as FE_TERN3

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _setup(_lib)
    return _lib


class OGeom(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("nz", C.c_int), ("g", C.c_int), ("gz", C.c_int),
                ("sy", C.c_long), ("sz", C.c_long), ("ncg", C.c_long),
                ("nfx", C.c_long), ("nfy", C.c_long), ("nfz", C.c_long),
                ("dx", C.c_double * 3), ("periodic", C.c_int)]


class EllipticOp(C.Structure):
    _fields_ = [("D", C.c_void_p), ("D_const", C.c_double),
                ("C", C.c_void_p), ("C_const", C.c_double),
                ("M", C.c_void_p), ("M_const", C.c_double),
                ("dx", C.c_double * 3)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_G = C.POINTER(OGeom)
_P = C.POINTER(ModelParams)


def _setup(L):
    sig = {
        "o_geom_init": (None, [_G, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
        "o_interp_func": (C.c_double, [C.c_double, C.c_char]),
        "o_deriv_interp_func": (C.c_double, [C.c_double, C.c_char]),
        "o_second_deriv_interp_func": (C.c_double, [C.c_double, C.c_char]),
        "o_well_func": (C.c_double, [C.c_double, C.c_char]),
        "o_deriv_well_func": (C.c_double, [C.c_double, C.c_char]),
        "o_second_deriv_well_func": (C.c_double, [C.c_double, C.c_char]),
        "o_average_func": (C.c_double, [C.c_double, C.c_double, C.c_char]),
        "o_eval_grad_normi": (C.c_double, [C.c_double, C.c_char, C.c_double, C.c_double]),
        "o_fenergy": (C.c_double, [_P, C.c_double, C.c_int]),
        "o_dfenergy": (C.c_double, [_P, C.c_double, C.c_int]),
        "o_d2fenergy": (C.c_double, [_P, C.c_double, C.c_int]),
        "o_kks_cell": (C.c_int, [_P, C.c_double, C.c_double, _dp, _dp]),
        "o_ghost_fill_periodic": (None, [_G, _dp, C.c_int]),
        "o_ghost_fill": (None, [_G, _dp, C.c_int]),
        "o_phase_flux": (None, [_G, _P, _dp, _dp, _dp]),
        "o_quat_grad_modulus": (None, [_G, _P, _dp, _dp]),
        "o_quat_face_norm2": (None, [_G, _P, _dp, _dp]),
        "o_phase_rhs": (None, [_G, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_quat_mobility": (None, [_G, _P, _dp, _dp]),
        "o_quat_face_coef": (None, [_G, _P, _dp, _dp, _dp]),
        "o_quat_rhs": (None, [_G, _P, _dp, _dp, _dp, _dp]),
        "o_quat_project": (None, [_G, _P, _dp, _dp, _dp]),
        "o_conc_coeffs": (None, [_G, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_conc_rhs": (None, [_G, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_kks_solve": (C.c_int, [_G, _P, _dp, _dp, _dp, _dp, _dp, _dp]),
        "o_grains_label": (C.c_int, [_G, _dp, C.c_double, C.c_int, _ip, _ip]),
        "o_ebs_coeffs": (None, [_G, _P, _dp, _dp, _dp]),
        "o_ebs_coeffs_mob": (None, [_G, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_atomic_mobility": (C.c_double, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
        "o_phase_diffusion_mobility": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
        "o_conc_rhs_ebs": (None, [_G, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_tag_cells": (None, [_G, _P, _dp, C.c_double, _ip]),
        "o_phase_fac_C": (None, [_G, _P, C.c_double, _dp, _dp, _dp]),
        "o_phase_mobility": (C.c_double, [_P, C.c_double, C.c_double]),
        "o_phase_mobility_field": (None, [_G, _P, _dp, _dp, _dp]),
        "o_fac_smooth": (C.c_double, [_G, C.POINTER(EllipticOp), _dp, _dp, C.c_int, C.c_double, _dp]),
        "o_fac_residual": (C.c_double, [_G, C.POINTER(EllipticOp), _dp, _dp, _dp, _dp]),
        "o_dot": (C.c_double, [_G, C.c_int, _dp, _dp]),
        "o_wrms": (C.c_double, [_G, C.c_int, _dp, _dp]),
        "o_sim_create": (C.c_void_p, [_P, C.POINTER(SolverParams), C.c_int, C.c_int, C.c_int]),
        "o_sim_destroy": (None, [C.c_void_p]),
        "o_sim_set_state": (None, [C.c_void_p, _dp, _dp, _dp]),
        "o_sim_get_state": (None, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
        "o_sim_advance": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Stats)]),
        "o_sim_set_periodic": (None, [C.c_void_p, C.c_int]),
        "o_conc_rhs_ebs_iso": (None, [_G, _P, _dp, _dp, _dp, _dp]),
        "o_phase_rhs_multiorder": (None, [_G, _P, C.c_int, _dp, _dp, _dp]),
        "o_project_phases": (None, [_G, C.c_int, _dp, _dp, _dp]),
        "o_conc_rhs_ebs_4th": (None, [_G, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_sim_set_temperature_ramp": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                              C.c_void_p]),
        "o_bf16": (C.c_float, [C.c_float]),
        "o_calphad_gfun": (C.c_double, [C.POINTER(CalphadGfun), C.c_double]),
        "o_calphad_at": (None, [C.POINTER(CalphadTdep), C.c_double, C.POINTER(CalphadPhase)]),
        "o_kks_newton": (C.c_int, [C.POINTER(CalphadPhase), C.c_double, C.c_double, C.c_double,
                                   _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_double]),
        "o_kks_batch": (C.c_long, [_P, C.POINTER(CalphadTdep), C.c_long, _dp, _dp, _dp, _dp, _dp,
                                   C.c_int, C.c_double, _dp, _dp, _ip]),
        "o_mlp_forward": (None, [C.c_int, C.c_int, C.c_int, _fp, C.c_long, _fp, _dp, C.c_int]),
        "o_mlp_loss_grad": (C.c_double, [C.c_int, C.c_int, C.c_int, _fp, C.c_long, _fp, _fp,
                                         _dp, C.c_int]),
        "o_hier_create": (C.c_void_p, [C.c_int, C.c_int, _ip, C.c_void_p, _ip, _dp, C.c_int]),
        "o_hier_destroy": (None, [C.c_void_p]),
        "o_hier_nlev": (C.c_int, [C.c_void_p]),
        "o_hier_npatch": (C.c_int, [C.c_void_p]),
        "o_hier_ncg": (C.c_long, [C.c_void_p]),
        "o_hier_nf": (C.c_long, [C.c_void_p]),
        "o_hier_offsets": (None, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "o_hier_geom": (_G, [C.c_void_p, C.c_int]),
        "o_hier_ghost_fill_level": (None, [C.c_void_p, C.c_int, _dp, C.c_int, C.c_int]),
        "o_hier_ghost_fill": (None, [C.c_void_p, _dp, C.c_int, C.c_int]),
        "o_hier_transfer": (C.c_int, [C.c_void_p, C.c_void_p, _dp, _dp, C.c_int]),
        "o_tern3_phase_conc": (None, [_P, _dp, _dp, _dp]),
        "o_kks_tern3": (None, [_G, _P, C.c_int, _dp, _dp, _dp]),
        "o_tern3_driving_force": (C.c_double, [_P, C.c_int, _dp, _dp, C.c_int]),
        "o_phase_rhs_tern3": (None, [_G, _P, C.c_int, _dp, _dp, _dp, _dp]),
        "o_ebs3_coeffs": (None, [_G, _P, C.c_int, _dp, _dp, _dp, _dp]),
        "o_conc_rhs_ebs3": (None, [_G, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_hier_weights": (None, [C.c_void_p, _dp]),
        "o_hier_coarsen_level": (None, [C.c_void_p, C.c_int, _dp, C.c_int, C.c_int]),
        "o_hier_coarsen": (None, [C.c_void_p, _dp, C.c_int, C.c_int]),
        "o_hier_coarsen_faces": (None, [C.c_void_p, C.c_int, _dp, C.c_int, C.c_long]),
        "o_hier_phase_rhs": (None, [C.c_void_p, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_hier_quat_face_norm2": (None, [C.c_void_p, _P, _dp, _dp]),
        "o_hier_quat_rhs": (None, [C.c_void_p, _P, _dp, _dp, _dp, _dp]),
        "o_hier_conc_coeffs": (None, [C.c_void_p, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_hier_conc_rhs": (None, [C.c_void_p, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_hier_phase_rhs_tern3": (None, [C.c_void_p, _P, C.c_int, _dp, _dp, _dp]),
    "o_hier_phase_rhs_multiorder": (None, [C.c_void_p, _P, C.c_int, _dp, _dp]),
        "o_hier_ebs3_coeffs": (None, [C.c_void_p, _P, C.c_int, _dp, _dp, _dp, _dp]),
        "o_hier_conc_rhs_ebs3": (None, [C.c_void_p, _P, _dp, _dp, _dp, _dp, _dp]),
        "o_sim_create_hier": (C.c_void_p, [_P, C.POINTER(SolverParams), C.c_void_p]),
        "o_sim_set_state_composite": (None, [C.c_void_p, _dp]),
        "o_sim_set_time": (None, [C.c_void_p, C.c_double, C.c_double]),
        "o_sim_get_state_composite": (None, [C.c_void_p, _dp]),
        "o_hier_comp_residual": (C.c_double, [C.c_void_p, C.c_void_p, _dp, _dp, _dp, C.c_int]),
        "o_hier_fac_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, C.c_int,
                                       C.c_double, _dp, _ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class Patch:
    """One periodic patch of the oracle layout (cell arrays with g ghosts)."""

    def __init__(self, ndim, n, g=1, dx=(1.0, 1.0, 1.0)):
        self.G = OGeom()
        nx, ny, nz = (n if isinstance(n, (tuple, list)) else (n,) * 3)
        dxa = (C.c_double * 3)(*dx)
        lib().o_geom_init(C.byref(self.G), ndim, nx, ny, nz if ndim == 3 else 1,
                          g, dxa)
        G = self.G
        self.ndim, self.nx, self.ny, self.nz, self.g = ndim, G.nx, G.ny, G.nz, g
        self.gz = G.gz
        self.ncg = G.ncg
        self.nf = G.nfx + G.nfy + G.nfz
        self.N = G.nx * G.ny * G.nz

    @property
    def gshape(self):
        """numpy shape of one ghosted cell array (k, j, i)."""
        return (self.nz + 2 * self.gz, self.ny + 2 * self.g, self.nx + 2 * self.g)

    def cells(self, ncomp=1):
        return np.zeros((ncomp,) + self.gshape)

    def sides(self, ncomp=1):
        return np.zeros((ncomp, self.nf))

    def interior(self, a):
        g, gz = self.g, self.gz
        return a[..., gz:gz + self.nz, g:g + self.ny, g:g + self.nx]

    def load(self, a, interior_vals):
        """interior_vals: (..., [nz,] ny, nx) -> ghosted array a (periodic
        ghost fill afterwards)."""
        v = np.asarray(interior_vals, dtype=np.float64)
        if self.ndim == 2 and (v.ndim < 3 or v.shape[-3] != 1 or
                               v.shape[-3:] != (1, self.ny, self.nx)):
            v = v.reshape(v.shape[:-2] + (1,) + v.shape[-2:])
        self.interior(a)[...] = v
        ncomp = int(np.prod(a.shape[:-3])) if a.ndim > 3 else 1
        lib().o_ghost_fill_periodic(C.byref(self.G), ptr(a), ncomp)
        return a


class Sim:
    """Oracle integrator handle."""

    def __init__(self, p, s, n, periodic=True):
        nx, ny, nz = (n if isinstance(n, (tuple, list)) else (n,) * 3)
        self.p, self.s = p, s
        self.h = lib().o_sim_create(C.byref(p), C.byref(s), nx, ny, nz)
        if not self.h:
            raise ValueError("o_sim_create rejected the model")
        self.shape = (ny, nx) if p.ndim == 2 else (nz, ny, nx)
        # phase orders and species (norderp > 1; the three-phase ternary model)
        self.nord = p.norderp if p.norderp > 1 else 1
        self.ncs = (2 if p.fe_model == FE_TERN3 else 1) if p.with_conc else 0
        if not periodic:
            lib().o_sim_set_periodic(self.h, 0)

    def set_temperature_ramp(self, T0, dTdt, T_target, tab=None):
        """T(t) = T0 + dTdt t held at T_target; CALPHAD phases from tab."""
        self._tab = tab
        lib().o_sim_set_temperature_ramp(self.h, T0, dTdt, T_target,
                                         C.cast(C.byref(tab), C.c_void_p) if tab is not None
                                         else None)

    def __del__(self):
        if getattr(self, "h", None):
            lib().o_sim_destroy(self.h)
            self.h = None

    def set_state(self, phi, c, q):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64) if c is not None else None
        q = np.ascontiguousarray(q, dtype=np.float64) if q is not None else None
        lib().o_sim_set_state(self.h, ptr(phi), ptr(c), ptr(q))

    def get_state(self):
        """phi, c, q, cl, cs; with several orders phi is (norderp, ...), with the
        three-phase ternary model c, cl (liquid) and cs (solid A) are (2, ...)."""
        lead = lambda k: (k,) if k > 1 else ()  # noqa: E731
        phi = np.zeros(lead(self.nord) + self.shape)
        c = np.zeros(lead(self.ncs) + self.shape)
        q = np.zeros((self.p.qlen,) + self.shape)
        cl = np.zeros(lead(self.ncs) + self.shape)
        cs = np.zeros(lead(self.ncs) + self.shape)
        lib().o_sim_get_state(self.h, ptr(phi), ptr(c), ptr(q), ptr(cl), ptr(cs))
        return phi, c, q, cl, cs

    def advance(self, nsteps):
        st = Stats()
        ret = lib().o_sim_advance(self.h, nsteps, C.byref(st))
        return ret, st.as_dict()


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def calphad_at(tab, T):
    out = (CalphadPhase * 2)()
    lib().o_calphad_at(C.byref(tab), T, out)
    return out


def kks_batch(p, tab, c, phi, T, gl=None, gs=None, probe_its=3, probe_tol=1.0e-8):
    """Per-cell-temperature CALPHAD KKS from (0.5, 0.5) or from (gl, gs)
    with the probe / restart rule; returns cl, cs, iters, failures."""
    c, phi, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (c, phi, T))
    n = c.size
    cl, cs = np.zeros(n), np.zeros(n)
    it = np.zeros(n, dtype=np.int32)
    if gl is not None:
        gl = np.ascontiguousarray(gl, dtype=np.float64)
        gs = np.ascontiguousarray(gs, dtype=np.float64)
    nf = lib().o_kks_batch(C.byref(p), C.byref(tab), n, ptr(c), ptr(phi), ptr(T), ptr(gl),
                           ptr(gs), probe_its, probe_tol, ptr(cl), ptr(cs), iptr(it))
    return cl, cs, it, nf


def mlp_forward(w, x, nin, hid, nout, emulate):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[-1]
    y = np.zeros((nout, n))
    lib().o_mlp_forward(nin, hid, nout, fptr(np.ascontiguousarray(w, dtype=np.float32)), n,
                        fptr(x), ptr(y), emulate)
    return y


def mlp_loss_grad(w, x, t, nin, hid, nout, emulate):
    w = np.ascontiguousarray(w, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(w.size)
    loss = lib().o_mlp_loss_grad(nin, hid, nout, fptr(w), x.shape[-1], fptr(x), fptr(t), ptr(g),
                                 emulate)
    return loss, g


# ---------------------------------------------------------------- hierarchy
class HierSim:
    """Oracle integrator on a hierarchy (composite state vectors)."""

    def __init__(self, p, s, hier):
        self.hier = hier
        self.p, self.s = p, s
        self.h = lib().o_sim_create_hier(C.byref(p), C.byref(s), hier.h)
        if not self.h:
            raise ValueError("o_sim_create_hier rejected the model")
        nord = p.norderp if p.norderp > 1 else 1
        ncs = (2 if p.fe_model == FE_TERN3 else 1) if p.with_conc else 0
        self.nc = nord + ncs + (p.qlen if p.evolve_quat else 0)

    def set_state(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        assert y.size == self.nc * self.hier.ncg
        lib().o_sim_set_state_composite(self.h, ptr(y))

    def get_state(self):
        y = np.zeros(self.nc * self.hier.ncg)
        lib().o_sim_get_state_composite(self.h, ptr(y))
        return y

    def set_time(self, t, h0=0.0):
        lib().o_sim_set_time(self.h, t, h0)

    def advance(self, nsteps):
        st = Stats()
        ret = lib().o_sim_advance(self.h, nsteps, C.byref(st))
        return ret, st.as_dict()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.o_sim_destroy(self.h)
            self.h = None


class HierBlock(C.Structure):
    """One preconditioner block of the hierarchy FAC (OHBlock)."""
    _fields_ = [("kind", C.c_int), ("D", C.c_void_p), ("D_const", C.c_double),
                ("C", C.c_void_p), ("C_const", C.c_double), ("M", C.c_void_p),
                ("M_const", C.c_double), ("S", C.c_void_p), ("gamma", C.c_double),
                ("ewing", C.c_int)]


class HierFacParams(C.Structure):
    _fields_ = [("max_cycles", C.c_int), ("pre", C.c_int), ("post", C.c_int),
                ("coarse_its", C.c_int), ("coarse_tol", C.c_double)]


class Hier:
    """Oracle hierarchy (OHier).  levels: list over levels of lists of
    (lo, hi) boxes, inclusive, in each level's index space; dom0: level-0
    domain cells (nx, ny[, nz])."""

    def __init__(self, ndim, levels, dom0, dx0=(1.0, 1.0, 1.0), periodic=True):
        from ampe_amd.params import Box
        self.ndim = ndim
        self.levels = [[(tuple(lo), tuple(hi)) for lo, hi in lev] for lev in levels]
        flat = [b for lev in self.levels for b in lev]
        arr = (Box * len(flat))()
        for i, (lo, hi) in enumerate(flat):
            arr[i].lo[:] = list(lo) + [0] * (3 - len(lo))
            arr[i].hi[:] = list(hi) + [0] * (3 - len(hi))
        npat = (C.c_int * len(levels))(*[len(v) for v in levels])
        d0 = (C.c_int * 3)(*(list(dom0) + [1] * (3 - len(dom0))))
        dx = (C.c_double * 3)(*dx0)
        self.h = lib().o_hier_create(ndim, len(levels), npat, arr, d0, dx, 1 if periodic else 0)
        if not self.h:
            raise ValueError("o_hier_create rejected the hierarchy")
        self.np = lib().o_hier_npatch(self.h)
        self.ncg = lib().o_hier_ncg(self.h)
        self.nf = lib().o_hier_nf(self.h)
        co = (C.c_long * self.np)()
        fo = (C.c_long * self.np)()
        lib().o_hier_offsets(self.h, co, fo)
        self.coff, self.foff = list(co), list(fo)
        self.geoms = [lib().o_hier_geom(self.h, p).contents for p in range(self.np)]
        self.patch_level = [l for l, lev in enumerate(self.levels) for _ in lev]
        self.boxes = flat
        self.dom0 = tuple(dom0)
        self.dx0 = tuple(dx0)
        self.periodic = bool(periodic)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.o_hier_destroy(self.h)
            self.h = None

    def cells(self, ncomp=1):
        return np.zeros(ncomp * self.ncg)

    def sides(self, ncomp=1):
        return np.zeros(ncomp * self.nf)

    def _blk(self, p):
        G = self.geoms[p]
        gz = G.gz
        return G, (G.nz + 2 * gz, G.ny + 2 * G.g, G.nx + 2 * G.g)

    def set_level(self, a, lev, vals, comp=0):
        """Level array vals (level domain, [nz,] ny, nx) -> interior cells of
        every patch of level lev in the composite array a."""
        v = np.asarray(vals, dtype=np.float64)
        v3 = v if self.ndim == 3 else v[None]
        for p in range(self.np):
            if self.patch_level[p] != lev:
                continue
            G, sh = self._blk(p)
            lo, hi = self.boxes[p]
            blk = a[comp * self.ncg + self.coff[p]:comp * self.ncg + self.coff[p] + G.ncg].reshape(sh)
            z0 = lo[2] if self.ndim == 3 else 0
            blk[G.gz:G.gz + G.nz, 1:1 + G.ny, 1:1 + G.nx] = \
                v3[z0:z0 + G.nz, lo[1]:lo[1] + G.ny, lo[0]:lo[0] + G.nx]

    def get_level(self, a, lev, comp=0, fill=np.nan):
        """Composite array -> level array over the level domain (cells off
        the level's patches = fill)."""
        n = [self.dom0[d] << lev for d in range(self.ndim)]
        shape = (n[2], n[1], n[0]) if self.ndim == 3 else (1, n[1], n[0])
        out = np.full(shape, fill)
        for p in range(self.np):
            if self.patch_level[p] != lev:
                continue
            G, sh = self._blk(p)
            lo, hi = self.boxes[p]
            blk = a[comp * self.ncg + self.coff[p]:comp * self.ncg + self.coff[p] + G.ncg].reshape(sh)
            z0 = lo[2] if self.ndim == 3 else 0
            out[z0:z0 + G.nz, lo[1]:lo[1] + G.ny, lo[0]:lo[0] + G.nx] = \
                blk[G.gz:G.gz + G.nz, 1:1 + G.ny, 1:1 + G.nx]
        return out if self.ndim == 3 else out[0]

    def ghost_view(self, a, p, comp=0):
        G, sh = self._blk(p)
        blk = a[comp * self.ncg + self.coff[p]:comp * self.ncg + self.coff[p] + G.ncg].reshape(sh)
        return blk if self.ndim == 3 else blk[0]

    def weights(self):
        w = self.cells()
        lib().o_hier_weights(self.h, ptr(w))
        return w

    def uncovered(self, lev):
        """Level-domain mask of the level's cells not under a finer level."""
        w = self.weights()
        m = self.get_level(w, lev, fill=0.0)
        return m > 0


def hier_transfer(Ho, Hn, fo, ncomp=1):
    """Interior cells of a field on Hn from fo on Ho (o_hier_transfer):
    copied from the old same level, else linear refinement of the new
    coarser level."""
    fo = np.ascontiguousarray(fo, dtype=np.float64)
    fn = Hn.cells(ncomp)
    if lib().o_hier_transfer(Ho.h, Hn.h, ptr(fo), ptr(fn), ncomp):
        raise ValueError("o_hier_transfer: hierarchies do not match or are not nested")
    return fn


def hier_level_tags(H, f, lev, p, tol):
    """Gradient-detector tags (o_tag_cells) of level lev over the level
    domain ([nz,] ny, nx), -1 off the level's patches; f: composite field
    with ghosts filled."""
    n = [H.dom0[d] << lev for d in range(H.ndim)]
    shape = tuple(n[::-1])
    tags = np.full(shape if H.ndim == 3 else (1,) + shape, -1, dtype=np.int32)
    for pp in range(H.np):
        if H.patch_level[pp] != lev:
            continue
        G = H.geoms[pp]
        blk = np.ascontiguousarray(f[H.coff[pp]:H.coff[pp] + G.ncg])
        t = np.zeros(G.nx * G.ny * G.nz, dtype=np.int32)
        lib().o_tag_cells(C.byref(G), C.byref(p), ptr(blk), tol, iptr(t))
        lo, _ = H.boxes[pp]
        z0 = lo[2] if H.ndim == 3 else 0
        tags[z0:z0 + G.nz, lo[1]:lo[1] + G.ny, lo[0]:lo[0] + G.nx] = t.reshape(G.nz, G.ny, G.nx)
    return tags if H.ndim == 3 else tags[0]


def regrid_levels(Ho, phi, p, tol, efficiency_tol, min_size, max_size, tag_buffer, max_levels):
    """New level boxes from the gradient detector, coarsest first (the role
    of GriddingAlgorithm::regridAllFinerLevels in PFModel::Regrid,
    source/PFModel.cc:565-590; tags of QuatModel::applyGradientDetector,
    source/QuatModel.cc:3437-3602): level 0 kept; for k = 0 .. max_levels-2
    the tag field phi (composite on Ho) is carried onto the levels 0..k
    built so far, its ghosts filled by linear refinement, level k tagged and
    clustered (cluster.cluster): over the level-0 domain for k = 0, above per
    patch of level k inside the patch shrunk by one cell on every side that
    is not a non-periodic physical boundary; the boxes refined by 2 are
    level k+1, and a level without tags ends the hierarchy."""
    from oracle import cluster as CL
    nd = Ho.ndim
    levels = [list(Ho.levels[0])]
    for k in range(max_levels - 1):
        Hk = Hier(nd, levels, Ho.dom0, Ho.dx0, Ho.periodic)
        f = hier_transfer(Ho, Hk, phi)
        lib().o_hier_ghost_fill(Hk.h, ptr(f), 1, 2)
        tags = hier_level_tags(Hk, f, k, p, tol)
        n = [Ho.dom0[d] << k for d in range(nd)]
        if k == 0:
            regions = [(tuple([0] * nd), tuple(x - 1 for x in n))]
        else:
            regions = []
            for (lo, hi) in levels[k]:
                lo, hi = list(lo), list(hi)
                for a in range(nd):
                    if Ho.periodic or lo[a] > 0:
                        lo[a] += 1
                    if Ho.periodic or hi[a] < n[a] - 1:
                        hi[a] -= 1
                if all(hi[a] >= lo[a] for a in range(nd)):
                    regions.append((tuple(lo), tuple(hi)))
        boxes = []
        for lo, hi in regions:
            sl = tuple(slice(lo[a], hi[a] + 1) for a in reversed(range(nd)))
            mask = (tags[sl] > 0).astype(np.uint8)
            for blo, bhi in CL.cluster(mask, efficiency_tol, min_size, max_size, tag_buffer):
                boxes.append((tuple(blo[a] + lo[a] for a in range(nd)),
                              tuple(bhi[a] + lo[a] for a in range(nd))))
        if not boxes:
            break
        levels.append([(tuple(2 * x for x in blo), tuple(2 * x + 1 for x in bhi))
                       for blo, bhi in boxes])
    return levels

