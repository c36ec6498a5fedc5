"""ctypes binding of libampe_hip.so (C ABI in include/ampe_hip.h).

There is no fallback: if the HIP library is missing or fails to load, every
entry point raises.  Device buffers are passed as raw pointers (torch
tensors' data_ptr() in the tests and the benchmark).
"""
import ctypes as C
import os

import numpy as np

from ampe_amd import params as P
from ampe_amd.params import (Box, CalphadPhase, CalphadTdep, KksBatchStats, ModelParams,
                             SolverParams, Stats, TrainParams)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libampe_hip.so")
_lib = None

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_ll = C.c_longlong


class HierBlock(C.Structure):
    """ampe_hier_block"""
    _fields_ = [("kind", _i), ("D", _vp), ("D_const", _d), ("C", _vp), ("C_const", _d),
                ("M", _vp), ("M_const", _d), ("S", _vp), ("gamma", _d), ("ewing", _i)]


class FacParams(C.Structure):
    """ampe_fac_params"""
    _fields_ = [("max_cycles", _i), ("pre", _i), ("post", _i), ("coarse_its", _i),
                ("coarse_tol", _d)]


class EllipticOp(C.Structure):
    _fields_ = [("D", _vp), ("D_const", _d), ("C", _vp), ("C_const", _d),
                ("M", _vp), ("M_const", _d), ("dx", _d * 3)]


_SIGS = {
    "ampe_last_error": (C.c_char_p, []),
    "ampe_version": (_i, []),
    "ampe_level_create": (_i, [_i, C.POINTER(Box), _i, _i, C.POINTER(Box), C.POINTER(_vp)]),
    "ampe_level_destroy": (_i, [_vp]),
    "ampe_level_sizes": (_i, [_vp, C.POINTER(_ll), C.POINTER(_ll), C.POINTER(_ll)]),
    "ampe_level_patch_offsets": (_i, [_vp, C.POINTER(_ll), C.POINTER(_ll)]),
    "ampe_level_patch_strides": (_i, [_vp, C.POINTER(_ll), C.POINTER(_ll)]),
    "ampe_sim_set_temperature_ramp": (_i, [_vp, _d, _d, _d, _vp]),
    "ampe_cluster_mask": (_i, [_i, _vp, _vp, _vp, _vp, _i, _vp]),
    "ampe_quat_coarsen": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "ampe_conc_rhs_ebs_iso": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp]),
    "ampe_linear_refine": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "ampe_phase_rhs_multiorder": (_i, [_vp, C.POINTER(ModelParams), _i, _vp, _vp, _vp, _vp]),
    "ampe_project_phases": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "ampe_conc_rhs_ebs_4th": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_quat_refine": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "ampe_cluster_tags": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp]),
    "ampe_ghost_fill": (_i, [_vp, _vp, _i, _vp]),
    "ampe_quat_grad": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp]),
    "ampe_phase_rhs": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_quat_rhs": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp]),
    "ampe_quat_project": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp]),
    "ampe_conc_coeffs": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_conc_rhs": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_kks_solve": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp, _vp,
                            C.POINTER(_i), _vp]),
    "ampe_quat_face_coefs": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp]),
    "ampe_quat_rhs_fc": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp]),
    "ampe_quat_smooth": (_i, [_vp, _vp, _vp, C.c_double, C.POINTER(C.c_double), C.c_int, _vp,
                              _vp, C.c_int, C.c_int, _vp]),
    "ampe_grains_label": (_i, [_vp, _vp, C.c_double, _vp, C.POINTER(C.c_int), _vp]),
    "ampe_kernel_timer_enable": (_i, [C.c_int, C.c_int]),
    "ampe_kernel_timer_read": (_i, [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                    C.POINTER(C.c_double)]),
    "ampe_conc_coeffs_ebs": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp]),
    "ampe_conc_rhs_ebs": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_tag_cells": (_i, [_vp, C.POINTER(ModelParams), _vp, _d, _vp, _vp]),
    "ampe_fac_phase_coeffs": (_i, [_vp, C.POINTER(ModelParams), _d, _vp, _vp, _vp, _vp]),
    "ampe_phase_mobility": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp]),
    "ampe_fac_smooth": (_i, [_vp, C.POINTER(EllipticOp), _vp, _vp, _i, _d, C.POINTER(_d), _vp]),
    "ampe_fac_residual": (_i, [_vp, C.POINTER(EllipticOp), _vp, _vp, _vp, C.POINTER(_d), _vp]),
    "ampe_vec_dot": (_i, [_vp, _i, _vp, _vp, C.POINTER(_d), _vp]),
    "ampe_vec_wrms": (_i, [_vp, _i, _vp, _vp, C.POINTER(_d), _vp]),
    "ampe_sim_create": (_i, [C.POINTER(ModelParams), C.POINTER(SolverParams), C.POINTER(Box),
                             _i, C.POINTER(Box), C.POINTER(_vp)]),
    "ampe_sim_destroy": (_i, [_vp]),
    "ampe_sim_set_state": (_i, [_vp, _vp, _vp, _vp]),
    "ampe_sim_get_state": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_sim_advance": (_i, [_vp, _i, C.POINTER(Stats)]),
    "ampe_sim_last_time_ms": (_i, [_vp, C.POINTER(_d)]),
    "ampe_sim_set_surrogate": (_i, [_vp, _vp, _i, _d, _d]),
    "ampe_coarsen_cell": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "ampe_coarsen_side": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "ampe_refine_add": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    "ampe_fac_residual_restrict": (_i, [_vp, _vp, C.POINTER(EllipticOp), _vp, _vp, _vp, _vp, _i,
                                        _vp]),
    "ampe_conc_coeffs_ebs_mobility": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_quat_residual_restrict": (_i, [_vp, _vp, _vp, _vp, C.c_double, _vp, _i, _vp, _vp,
                                         _vp, _vp, _i, _vp]),
    "ampe_calphad_at": (_i, [C.POINTER(CalphadTdep), _d, C.POINTER(CalphadPhase)]),
    "ampe_surrogate_create": (_i, [_i, _i, _i, C.c_ulonglong, C.POINTER(_vp)]),
    "ampe_surrogate_destroy": (_i, [_vp]),
    "ampe_surrogate_nparams": (_i, [_vp, C.POINTER(_ll)]),
    "ampe_surrogate_set_params": (_i, [_vp, _vp]),
    "ampe_surrogate_get_params": (_i, [_vp, _vp]),
    "ampe_surrogate_set_scaling": (_i, [_vp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d),
                                        C.POINTER(_d)]),
    "ampe_surrogate_forward": (_i, [_vp, _ll, _vp, _vp, _vp]),
    "ampe_surrogate_loss_grad": (_i, [_vp, _ll, _vp, _vp, _vp, C.POINTER(_d), _vp]),
    "ampe_surrogate_train": (_i, [_vp, _ll, _vp, _vp, C.POINTER(TrainParams), _vp, _vp]),
    "ampe_kks_batch": (_i, [C.POINTER(ModelParams), C.POINTER(CalphadTdep), _ll, _vp, _vp, _vp,
                            _vp, _i, _d, _vp, _vp, _vp, C.POINTER(KksBatchStats), _vp]),
    "ampe_kks_batch_noalloc": (_i, [C.POINTER(ModelParams), C.POINTER(CalphadTdep), _ll, _vp,
                                    _vp, _vp, _vp, _i, _d, _vp, _vp, _vp, _vp, _vp]),
    "ampe_sim_vector_size": (_i, [_vp, C.POINTER(_ll), C.POINTER(_i)]),
    "ampe_sim_create_hier": (_i, [C.POINTER(ModelParams), C.POINTER(SolverParams), _vp,
                                  C.POINTER(_vp)]),
    "ampe_sim_set_state_composite": (_i, [_vp, _vp]),
    "ampe_sim_get_state_composite": (_i, [_vp, _vp]),
    "ampe_hier_create": (_i, [_i, _i, _vp, _vp, _vp, _vp, _i, C.POINTER(_vp)]),
    "ampe_hier_boxes": (_i, [_vp, C.POINTER(_i), _vp, _vp, _i]),
    "ampe_kks_ternary3": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp]),
    "ampe_phase_rhs_ternary3": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp]),
    "ampe_ebs3_coeffs": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp]),
    "ampe_conc_rhs_ebs3": (_i, [_vp, C.POINTER(ModelParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_hier_transfer": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "ampe_hier_regrid": (_i, [_vp, _vp, C.POINTER(ModelParams), _d, _vp, _i, C.POINTER(_vp),
                              _vp]),
    "ampe_sim_set_time": (_i, [_vp, _d, _d]),
    "ampe_hier_destroy": (_i, [_vp]),
    "ampe_hier_sizes": (_i, [_vp, C.POINTER(_ll), C.POINTER(_ll), C.POINTER(_ll)]),
    "ampe_hier_level": (_i, [_vp, _i, C.POINTER(_vp)]),
    "ampe_hier_composite": (_i, [_vp, C.POINTER(_vp)]),
    "ampe_hier_weights": (_i, [_vp, _vp, _vp]),
    "ampe_hier_ghost_fill": (_i, [_vp, _vp, _i, _i, _vp]),
    "ampe_hier_coarsen": (_i, [_vp, _vp, _i, _i, _vp]),
    "ampe_hier_phase_rhs": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_hier_quat_face_norm2": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "ampe_hier_quat_rhs": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_hier_conc_coeffs": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_hier_conc_rhs": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ampe_hier_fac_solve": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _d, C.POINTER(_d), C.POINTER(_i),
                                 _vp]),
    "ampe_sim_level": (_i, [_vp, C.POINTER(_vp)]),
    "ampe_sim_stream": (_i, [_vp, C.POINTER(_vp)]),
    "ampe_rhs_eval": (_i, [_vp, _d, _vp, _vp, _i]),
    "ampe_precond_setup": (_i, [_vp, _d, _vp, _d, C.POINTER(_i)]),
    "ampe_precond_solve": (_i, [_vp, _d, _vp, _vp, _d, _d]),
    "ampe_apply_projection": (_i, [_vp, _vp, _vp, _vp]),
    "ampe_sim_reset_ref_concentrations": (_i, [_vp]),
    "ampe_sim_kks_status": (_i, [_vp, C.POINTER(Stats)]),
    "ampe_vec_axpby": (_i, [_ll, _d, _vp, _d, _vp, _vp, _vp]),
}

EXPORTED = tuple(_SIGS)


class AmpeError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AmpeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc < 0:
        raise AmpeError(f"{what}: {lib().ampe_last_error().decode()}")
    return rc


def boxes_of(shape, ndim, origin=(0, 0, 0)):
    """Single box covering an array of shape ([nz,] ny, nx)."""
    b = Box()
    n = list(shape[::-1]) + [1] * (3 - len(shape))
    for a in range(3):
        b.lo[a] = origin[a]
        b.hi[a] = origin[a] + n[a] - 1
    return b


class Level:
    """A patch level on the device (ampe_level)."""

    def __init__(self, ndim, boxes, ghost=1, periodic=None):
        self.ndim = ndim
        if isinstance(boxes, Box):
            boxes = [boxes]
        self.boxes = list(boxes)
        arr = (Box * len(self.boxes))(*self.boxes)
        h = _vp()
        per = C.byref(periodic) if periodic is not None else None
        check(lib().ampe_level_create(ndim, arr, len(self.boxes), ghost, per, C.byref(h)),
              "ampe_level_create")
        self.h = h
        ncg, nc = _ll(), _ll()
        ns = (_ll * 3)()
        check(lib().ampe_level_sizes(h, C.byref(ncg), C.byref(nc), ns), "ampe_level_sizes")
        self.ncg, self.ncell = ncg.value, nc.value
        self.nside = [ns[0], ns[1], ns[2]]
        self.nside_tot = sum(self.nside)
        np_ = len(self.boxes)
        co = (_ll * np_)()
        so = (_ll * (3 * np_))()
        check(lib().ampe_level_patch_offsets(h, co, so), "ampe_level_patch_offsets")
        self.cell_off = list(co)
        self.side_off = [list(so[3 * p:3 * p + 3]) for p in range(np_)]
        sy, sz = (_ll * np_)(), (_ll * np_)()
        check(lib().ampe_level_patch_strides(h, sy, sz), "ampe_level_patch_strides")
        self.sy, self.sz = list(sy), list(sz)
        self.ghost = ghost

    def ghost_view(self, buf, p, ghost=None):
        """View of patch p's ghosted block in the flat array buf (one
        component): shape ([nz + 2g,] ny + 2g, nx + 2g), row padding dropped."""
        g = self.ghost if ghost is None else ghost
        gz = g if self.ndim == 3 else 0
        n = self._patch_ext(self.boxes[p])
        rows = (n[2] + 2 * gz) * (n[1] + 2 * g)
        blk = buf[self.cell_off[p]:self.cell_off[p] + rows * self.sy[p]]
        blk = blk.reshape(n[2] + 2 * gz, n[1] + 2 * g, self.sy[p])[:, :, :n[0] + 2 * g]
        return blk if self.ndim == 3 else blk[0]

    def _patch_ext(self, b):
        n = [b.hi[a] - b.lo[a] + 1 for a in range(3)]
        if self.ndim == 2:
            n[2] = 1
        return n

    def to_ghosted(self, a, ghost=1):
        """Domain array ([nz,] ny, nx), domain corner at the patches' lowest
        corner, to the level's ghosted cell layout (ghost entries zero)."""
        a = np.asarray(a, dtype=np.float64)
        a3 = a if self.ndim == 3 else a[None]
        org = [min(b.lo[d] for b in self.boxes) for d in range(3)]
        out = np.zeros(self.ncg)
        g, gz = ghost, (ghost if self.ndim == 3 else 0)
        for p, b in enumerate(self.boxes):
            n = self._patch_ext(b)
            x0, y0, z0 = (b.lo[d] - org[d] for d in range(3))
            blk = self.ghost_view(out, p, ghost)
            blk = blk if self.ndim == 3 else blk[None]
            blk[gz:gz + n[2], g:g + n[1], g:g + n[0]] = \
                a3[z0:z0 + n[2], y0:y0 + n[1], x0:x0 + n[0]]
        return out

    def from_interior(self, v):
        """Interior-compact per-patch array (tags, grain labels) to the domain
        array ([nz,] ny, nx)."""
        org = [min(b.lo[d] for b in self.boxes) for d in range(3)]
        top = [max(b.hi[d] for b in self.boxes) for d in range(3)]
        shape = [top[d] - org[d] + 1 for d in range(3)]
        if self.ndim == 2:
            shape[2] = 1
        out = np.zeros((shape[2], shape[1], shape[0]), dtype=v.dtype)
        off = 0
        for b in self.boxes:
            n = self._patch_ext(b)
            x0, y0, z0 = (b.lo[d] - org[d] for d in range(3))
            cnt = n[0] * n[1] * n[2]
            out[z0:z0 + n[2], y0:y0 + n[1], x0:x0 + n[0]] = \
                v[off:off + cnt].reshape(n[2], n[1], n[0])
            off += cnt
        return out if self.ndim == 3 else out[0]

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.ampe_level_destroy(self.h)
            self.h = None


class Sim:
    """Device-resident integrator (ampe_sim)."""

    def __init__(self, p, s, boxes, periodic=None):
        if isinstance(boxes, Box):
            boxes = [boxes]
        arr = (Box * len(boxes))(*boxes)
        h = _vp()
        per = C.byref(periodic) if periodic is not None else None
        check(lib().ampe_sim_create(C.byref(p), C.byref(s), arr, len(boxes), per, C.byref(h)),
              "ampe_sim_create")
        self.h = h
        self.p, self.s = p, s
        self.ncell = sum((b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) *
                         ((b.hi[2] - b.lo[2] + 1) if p.ndim == 3 else 1) for b in boxes)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.ampe_sim_destroy(self.h)
            self.h = None

    @staticmethod
    def _ptr(a):
        if a is None:
            return None
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return a.ctypes.data_as(_vp)

    def set_time(self, t, h0=0.0):
        """Restart at time t with first step h0 (ampe_sim_set_time)."""
        check(lib().ampe_sim_set_time(self.h, t, h0), "ampe_sim_set_time")

    def set_state(self, phi, c, q):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        c = None if c is None else np.ascontiguousarray(c, dtype=np.float64)
        q = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
        check(lib().ampe_sim_set_state(self.h, self._ptr(phi), self._ptr(c), self._ptr(q)),
              "ampe_sim_set_state")

    def get_state(self, shape):
        """phi, c, q, cl, cs; with several orders phi is (norderp, *shape), with
        the three-phase ternary model c, cl (liquid) and cs (solid A) are
        (2, *shape)."""
        shape = tuple(shape)
        nord = self.p.norderp if self.p.norderp > 1 else 1
        ncs = 2 if self.p.fe_model == P.FE_QUADRATIC_TERNARY_3PH else 1
        lead = lambda k: (k,) if k > 1 else ()  # noqa: E731
        phi = np.zeros(lead(nord) + shape)
        c = np.zeros(lead(ncs) + shape)
        q = np.zeros((self.p.qlen,) + shape)
        cl = np.zeros(lead(ncs) + shape)
        cs = np.zeros(lead(ncs) + shape)
        check(lib().ampe_sim_get_state(self.h, self._ptr(phi), self._ptr(c), self._ptr(q),
                                       self._ptr(cl), self._ptr(cs)), "ampe_sim_get_state")
        return phi, c, q, cl, cs

    def set_temperature_ramp(self, T0, dTdt, T_target, tab=None):
        """T(t) = T0 + dTdt t, held at T_target; CALPHAD phases re-evaluated
        from the CalphadTdep table tab at every step."""
        self._tab = tab
        check(lib().ampe_sim_set_temperature_ramp(
            self.h, T0, dTdt, T_target,
            C.cast(C.byref(tab), _vp) if tab is not None else None),
            "ampe_sim_set_temperature_ramp")

    def advance(self, nsteps):
        st = Stats()
        rc = check(lib().ampe_sim_advance(self.h, nsteps, C.byref(st)), "ampe_sim_advance")
        return rc, st.as_dict()

    def last_time_ms(self):
        ms = _d()
        check(lib().ampe_sim_last_time_ms(self.h, C.byref(ms)), "ampe_sim_last_time_ms")
        return ms.value

    def set_surrogate(self, surrogate, probe_its=3, probe_tol=1.0e-8, switch_tol=1.0e-2):
        h = surrogate.h if surrogate is not None else None
        check(lib().ampe_sim_set_surrogate(self.h, h, probe_its, probe_tol, switch_tol),
              "ampe_sim_set_surrogate")
        self._surrogate = surrogate  # keep it alive as long as the sim


def _dev_ptr(t):
    return None if t is None else _vp(t.data_ptr())


def _stream_ptr(stream):
    return None if stream is None else _vp(stream.cuda_stream)


class Surrogate:
    """KKS surrogate MLP (ampe_surrogate): n_in -> 64 -> 64 -> n_out, tanh,
    BF16 matrix-core kernels with FP32 accumulation.  Device arrays are torch
    tensors (float32 for normalised data, float64 for raw training data)."""

    def __init__(self, n_in=3, n_hidden=64, n_out=2, seed=0):
        h = _vp()
        check(lib().ampe_surrogate_create(n_in, n_hidden, n_out, seed, C.byref(h)),
              "ampe_surrogate_create")
        self.h = h
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        n = _ll()
        check(lib().ampe_surrogate_nparams(h, C.byref(n)), "ampe_surrogate_nparams")
        self.nparams = n.value

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.ampe_surrogate_destroy(self.h)
            self.h = None

    def get_params(self):
        w = np.zeros(self.nparams, dtype=np.float32)
        check(lib().ampe_surrogate_get_params(self.h, w.ctypes.data_as(_vp)),
              "ampe_surrogate_get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nparams
        check(lib().ampe_surrogate_set_params(self.h, w.ctypes.data_as(_vp)),
              "ampe_surrogate_set_params")

    def set_scaling(self, in_shift, in_scale, out_shift, out_scale):
        arr = lambda v: (_d * 4)(*([float(x) for x in v] + [0.0] * (4 - len(v))))  # noqa: E731
        self.scaling = (list(in_shift), list(in_scale), list(out_shift), list(out_scale))
        check(lib().ampe_surrogate_set_scaling(self.h, arr(in_shift), arr(in_scale),
                                               arr(out_shift), arr(out_scale)),
              "ampe_surrogate_set_scaling")

    def forward(self, x, y, stream=None):
        n = x.shape[-1]
        check(lib().ampe_surrogate_forward(self.h, n, _dev_ptr(x), _dev_ptr(y),
                                           _stream_ptr(stream)), "ampe_surrogate_forward")

    def loss_grad(self, x, t, grad, stream=None):
        loss = _d()
        check(lib().ampe_surrogate_loss_grad(self.h, x.shape[-1], _dev_ptr(x), _dev_ptr(t),
                                             _dev_ptr(grad), C.byref(loss),
                                             _stream_ptr(stream)), "ampe_surrogate_loss_grad")
        return loss.value

    def train(self, x, t, batch, steps, lr0=3e-3, lr1=3e-5, beta1=0.9, beta2=0.999,
              eps=1e-8, stream=None):
        tp = TrainParams(batch=batch, steps=steps, lr0=lr0, lr1=lr1, beta1=beta1,
                         beta2=beta2, eps=eps)
        hist = np.zeros(steps, dtype=np.float32)
        check(lib().ampe_surrogate_train(self.h, x.shape[-1], _dev_ptr(x), _dev_ptr(t),
                                         C.byref(tp), hist.ctypes.data_as(_vp),
                                         _stream_ptr(stream)), "ampe_surrogate_train")
        return hist


def kks_batch(p, tab, c, phi, T, cl, cs, iters=None, guess=None, probe_its=3,
              probe_tol=1.0e-8, stream=None):
    """Binary CALPHAD KKS over independent cells (ampe_kks_batch); c, phi, T,
    cl, cs: float64 device tensors, iters: int32 device tensor or None."""
    st = KksBatchStats()
    check(lib().ampe_kks_batch(C.byref(p), C.byref(tab), c.numel(), _dev_ptr(c), _dev_ptr(phi),
                               _dev_ptr(T), guess.h if guess is not None else None, probe_its,
                               probe_tol, _dev_ptr(cl), _dev_ptr(cs), _dev_ptr(iters),
                               C.byref(st), _stream_ptr(stream)), "ampe_kks_batch")
    return st.as_dict()


def kks_batch_noalloc(p, tab, c, phi, T, cl, cs, stats_dev, iters=None, guess=None,
                      probe_its=3, probe_tol=1.0e-8, stream=None):
    """ampe_kks_batch_noalloc: as kks_batch, statistics written to stats_dev
    (int64 device tensor of 4: updates, largest count, fallbacks, failures,
    or None); nothing is synchronised."""
    check(lib().ampe_kks_batch_noalloc(C.byref(p), C.byref(tab), c.numel(), _dev_ptr(c),
                                       _dev_ptr(phi), _dev_ptr(T),
                                       guess.h if guess is not None else None, probe_its,
                                       probe_tol, _dev_ptr(cl), _dev_ptr(cs), _dev_ptr(iters),
                                       _dev_ptr(stats_dev), _stream_ptr(stream)),
          "ampe_kks_batch_noalloc")


def cluster_mask(mask, cp, lo=(0, 0, 0), max_boxes=4096):
    """Boxes (list of (lo, hi) inclusive, 3 ints each) clustering the nonzero
    cells of mask ([nz,] ny, nx) whose first cell sits at lo (ampe_cluster_mask)."""
    m = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.uint8)
    ndim = m.ndim
    shape3 = m.shape[::-1] + (1,) * (3 - ndim)
    dom = Box()
    for d in range(3):
        dom.lo[d] = lo[d] if d < ndim else 0
        dom.hi[d] = dom.lo[d] + shape3[d] - 1
    out = (Box * max_boxes)()
    n = C.c_int(0)
    check(lib().ampe_cluster_mask(ndim, C.byref(dom), m.ctypes.data_as(_vp), C.byref(cp), out,
                                  max_boxes, C.byref(n)), "ampe_cluster_mask")
    return [(tuple(out[b].lo), tuple(out[b].hi)) for b in range(n.value)]


class Hier:
    """A patch hierarchy on the device (ampe_hier).  levels: list over levels
    of lists of (lo, hi) inclusive boxes in each level's index space; dom0:
    level-0 cells (nx, ny[, nz])."""

    def __init__(self, ndim, levels, dom0, dx0=(1.0, 1.0, 1.0), periodic=True, _handle=None):
        self.ndim = ndim
        self.dom0, self.dx0, self.periodic = tuple(dom0), tuple(dx0), bool(periodic)
        if _handle is not None:
            h = _handle
            levels = self._levels_of(h, ndim)
        self.levels = [[(tuple(lo), tuple(hi)) for lo, hi in lev] for lev in levels]
        flat = [b for lev in self.levels for b in lev]
        if _handle is None:
            arr = (Box * len(flat))()
            for i, (lo, hi) in enumerate(flat):
                arr[i].lo[:] = list(lo) + [0] * (3 - len(lo))
                arr[i].hi[:] = list(hi) + [0] * (3 - len(hi))
            npat = (_i * len(levels))(*[len(v) for v in levels])
            d0 = (_i * 3)(*(list(dom0) + [1] * (3 - len(dom0))))
            dx = (_d * 3)(*dx0)
            h = _vp()
            check(lib().ampe_hier_create(ndim, len(levels), npat, arr, d0, dx,
                                         1 if periodic else 0, C.byref(h)), "ampe_hier_create")
        self.h = h
        ncg, nc = _ll(), _ll()
        ns = (_ll * 3)()
        check(lib().ampe_hier_sizes(h, C.byref(ncg), C.byref(nc), ns), "ampe_hier_sizes")
        self.ncg, self.ncell = ncg.value, nc.value
        self.nside = [ns[0], ns[1], ns[2]]
        self.nside_tot = sum(self.nside)
        self.boxes = flat
        self.patch_level = [l for l, lev in enumerate(self.levels) for _ in lev]

    def patch_blocks(self, y, comp=0):
        """[(level, lo, hi, view)] per patch, in the composite order (levels,
        then patches): view is the patch's interior cells ([nz,] ny, nx) of
        component comp in the host composite array y, each patch a ghosted
        block with one ghost cell per side along each direction of the
        level."""
        out, off = [], 0
        for lev, (lo, hi) in zip(self.patch_level, self.boxes):
            n = [hi[d] - lo[d] + 1 for d in range(self.ndim)]
            sh = tuple(m + 2 for m in reversed(n))
            size = int(np.prod(sh))
            blk = y[comp * self.ncg + off:comp * self.ncg + off + size].reshape(sh)
            out.append((lev, lo, hi, blk[(slice(1, -1),) * self.ndim]))
            off += size
        assert off == self.ncg, (off, self.ncg)
        return out

    @staticmethod
    def _levels_of(h, ndim):
        nlev = _i()
        npat = (_i * 4)()
        check(lib().ampe_hier_boxes(h, C.byref(nlev), npat, None, 0), "ampe_hier_boxes")
        nb = sum(npat[l] for l in range(nlev.value))
        arr = (Box * nb)()
        check(lib().ampe_hier_boxes(h, C.byref(nlev), npat, arr, nb), "ampe_hier_boxes")
        out, b = [], 0
        for l in range(nlev.value):
            lev = []
            for _ in range(npat[l]):
                lev.append((tuple(arr[b].lo[:ndim]), tuple(arr[b].hi[:ndim])))
                b += 1
            out.append(lev)
        return out

    def regrid(self, phi_ptr, p0, tol, cp, max_levels, stream=None):
        """A new hierarchy from the gradient detector on the composite tag
        field at device address phi_ptr (ampe_hier_regrid)."""
        h = _vp()
        check(lib().ampe_hier_regrid(self.h, phi_ptr, C.byref(p0), tol, C.byref(cp), max_levels,
                                     C.byref(h), stream), "ampe_hier_regrid")
        return Hier(self.ndim, None, self.dom0, self.dx0, self.periodic, _handle=h)

    def transfer_to(self, new, f_old_ptr, f_new_ptr, ncomp, stream=None):
        """Interior cells of a field on the new hierarchy (ampe_hier_transfer)."""
        check(lib().ampe_hier_transfer(self.h, new.h, f_old_ptr, f_new_ptr, ncomp, stream),
              "ampe_hier_transfer")

    def level(self, l):
        v = _vp()
        check(lib().ampe_hier_level(self.h, l, C.byref(v)), "ampe_hier_level")
        return v

    def composite(self):
        v = _vp()
        check(lib().ampe_hier_composite(self.h, C.byref(v)), "ampe_hier_composite")
        return v

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.ampe_hier_destroy(self.h)
            self.h = None


class HierSim:
    """Device integrator on a hierarchy (ampe_sim_create_hier); states are
    host composite arrays (nc x the hierarchy's ghosted cells)."""

    def __init__(self, p, s, hier):
        h = _vp()
        check(lib().ampe_sim_create_hier(C.byref(p), C.byref(s), hier.h, C.byref(h)),
              "ampe_sim_create_hier")
        self.h = h
        self.hier = hier  # outlives the sim
        self.p, self.s = p, s
        nord = p.norderp if p.norderp > 1 else 1
        ncs = (2 if p.fe_model == P.FE_QUADRATIC_TERNARY_3PH else 1) if p.with_conc else 0
        self.nc = nord + ncs + (p.qlen if p.evolve_quat else 0)

    def set_state(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        assert y.size == self.nc * self.hier.ncg
        check(lib().ampe_sim_set_state_composite(self.h, y.ctypes.data_as(_vp)),
              "ampe_sim_set_state_composite")

    def get_state(self):
        y = np.zeros(self.nc * self.hier.ncg)
        check(lib().ampe_sim_get_state_composite(self.h, y.ctypes.data_as(_vp)),
              "ampe_sim_get_state_composite")
        return y

    def set_time(self, t, h0=0.0):
        check(lib().ampe_sim_set_time(self.h, t, h0), "ampe_sim_set_time")

    def advance(self, nsteps):
        st = Stats()
        rc = check(lib().ampe_sim_advance(self.h, nsteps, C.byref(st)), "ampe_sim_advance")
        return rc, st.as_dict()

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.ampe_sim_destroy(self.h)
            self.h = None
