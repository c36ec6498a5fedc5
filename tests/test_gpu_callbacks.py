"""The integrator callbacks driven from a host-side BDF / Newton / SPGMR loop,
as a reference host that keeps its own CVODE would drive them
(CVODEAbstractFunctions, source/samrai/CVODEAbstractFunctions.h:61-88):
ampe_rhs_eval, ampe_precond_setup / ampe_precond_solve, ampe_apply_projection,
ampe_vec_axpby and ampe_vec_dot on caller-owned device vectors.  The host loop
is the same algorithm as ampe_sim_advance (imposed step, BDF1 then
variable-step BDF2, CVODE's Newton tests, SPGMR with modified Gram-Schmidt
and left preconditioning); it must reproduce ampe_sim_advance's Newton and
Krylov counts and fields step by step, with the red-black preconditioner in
FP64 and with the FP32 multigrid preconditioner.
"""
import ctypes as C
import math

import numpy as np
import pytest

from ampe_amd import params as P

pytestmark = pytest.mark.gpu


class HostBDF:
    """Host loop over the library's callbacks (vectors: torch float64 tensors
    on the sim's stream)."""

    def __init__(self, H, sim, p, s, lev_layout, dev):
        import torch
        self.H, self.sim, self.p, self.s = H, sim, p, s
        n, nc = C.c_longlong(), C.c_int()
        H.check(H.lib().ampe_sim_vector_size(sim.h, C.byref(n), C.byref(nc)))
        self.nv, self.nc = n.value, nc.value
        lv = C.c_void_p()
        H.check(H.lib().ampe_sim_level(sim.h, C.byref(lv)))
        self.lev = lv
        sp = C.c_void_p()
        H.check(H.lib().ampe_sim_stream(sim.h, C.byref(sp)))
        self.sp = sp
        self.stream = torch.cuda.ExternalStream(sp.value)
        self.layout = lev_layout
        self.ncell = lev_layout.ncell
        self.cv = p.dx[0] * p.dx[1] * (p.dx[2] if p.ndim == 3 else 1.0)
        self.dev = dev
        self.t, self.h, self.hprev, self.nst, self.nstlp = 0.0, s.dt, 0.0, 0, 0
        self.gamma_setup, self.crate = 1.0, 1.0
        self.have_prev = False
        self.nni = self.nli = self.ncfn = 0

    def zeros(self):
        import torch
        with torch.cuda.stream(self.stream):
            return torch.zeros(self.nv, dtype=torch.float64, device=self.dev)

    def _p(self, t):
        return C.c_void_p(t.data_ptr())

    def dot(self, x, y):
        out = C.c_double()
        self.H.check(self.H.lib().ampe_vec_dot(self.lev, self.nc, self._p(x), self._p(y),
                                               C.byref(out), self.sp))
        return out.value * self.cv

    def wrms(self, x):
        out = C.c_double()
        self.H.check(self.H.lib().ampe_vec_wrms(self.lev, self.nc, self._p(x),
                                                self._p(self.ewt), C.byref(out), self.sp))
        return out.value

    def axpby(self, a, x, b, y, z):
        self.H.check(self.H.lib().ampe_vec_axpby(self.nv, a, self._p(x), b, self._p(y),
                                                 self._p(z), self.sp))

    def rhs(self, y, ydot, fd):
        self.H.check(self.H.lib().ampe_rhs_eval(self.sim.h, self.t + self.h, self._p(y),
                                                self._p(ydot), fd))

    def psolve(self, r, z, delta):
        return self.H.check(self.H.lib().ampe_precond_solve(self.sim.h, self.t, self._p(r),
                                                            self._p(z), self.gamma_setup, delta))

    def set_state(self, y):
        import torch
        self.y = y
        with torch.cuda.stream(self.stream):
            self.ewt = 1.0 / (self.s.rtol * torch.abs(y) + self.s.atol)

    def spgmr(self, gamma, b, x, delta):
        import torch
        st = self.stream
        maxl = self.s.maxl
        work = self.zeros()
        self.psolve(b, work, delta)
        with torch.cuda.stream(st):
            V = [self.ewt * work]
        rho = math.sqrt(self.dot(V[0], V[0]))
        rho0 = rho
        if rho <= delta:
            with torch.cuda.stream(st):
                x.zero_()
            return 0
        with torch.cuda.stream(st):
            V[0] = V[0] * (1.0 / rho)
        Hm = np.zeros((maxl + 1, maxl + 1))
        g = np.zeros(maxl + 1)
        g[0] = rho
        giv = []
        lused, converged = 0, False
        for l in range(maxl):
            self.nli += 1
            with torch.cuda.stream(st):
                xk = V[l] / self.ewt
            sig = 1.0 / self.wrms(xk)
            w2, jv = self.zeros(), self.zeros()
            self.axpby(sig, xk, 1.0, self.ycur, w2)
            self.rhs(w2, jv, 1)
            with torch.cuda.stream(st):
                jvn = (jv - self.fcur) * (1.0 / sig)
                v = xk - gamma * jvn
            self.psolve(v, work, delta)
            with torch.cuda.stream(st):
                v = self.ewt * work
            for i in range(l + 1):
                hij = self.dot(V[i], v)
                Hm[i, l] = hij
                self.axpby(1.0, v, -hij, V[i], v)
            hn = math.sqrt(self.dot(v, v))
            Hm[l + 1, l] = hn
            for i in range(l):
                c, s_ = giv[i]
                t1, t2 = Hm[i, l], Hm[i + 1, l]
                Hm[i, l] = c * t1 - s_ * t2
                Hm[i + 1, l] = s_ * t1 + c * t2
            a, bb = Hm[l, l], Hm[l + 1, l]
            if bb == 0.0:
                c, s_ = 1.0, 0.0
            elif abs(bb) > abs(a):
                tt = a / bb
                s_ = -1.0 / math.sqrt(1.0 + tt * tt)
                c = -s_ * tt
            else:
                tt = bb / a
                c = 1.0 / math.sqrt(1.0 + tt * tt)
                s_ = -c * tt
            giv.append((c, s_))
            Hm[l, l] = c * a - s_ * bb
            Hm[l + 1, l] = 0.0
            g[l + 1] = s_ * g[l]
            g[l] = c * g[l]
            rho = abs(g[l + 1])
            lused = l + 1
            if rho <= delta:
                converged = True
                break
            if hn == 0.0:
                break
            with torch.cuda.stream(st):
                V.append(v * (1.0 / hn))
        yv = np.zeros(lused)
        for i in range(lused - 1, -1, -1):
            yv[i] = (g[i] - sum(Hm[i, j] * yv[j] for j in range(i + 1, lused))) / Hm[i, i]
        with torch.cuda.stream(st):
            acc = torch.zeros_like(x)
            for i in range(lused):
                acc = acc + yv[i] * V[i]
            x.copy_(acc / self.ewt)
        if converged:
            return 0
        return 1 if rho < rho0 else 2

    def newton(self, gamma, delta, force_setup):
        import torch
        st = self.stream
        s = self.s
        call_setup = (force_setup or self.nst == 0 or abs(gamma / self.gamma_setup - 1.0) > 0.3
                      or self.nst >= self.nstlp + s.msbp)
        did_setup = False
        with torch.cuda.stream(st):
            self.ycur = self.ypred.clone()
        self.fcur = self.zeros()
        b, dl = self.zeros(), self.zeros()
        delp = 0.0
        for m in range(s.max_newton):
            self.nni += 1
            self.rhs(self.ycur, self.fcur, 0)
            with torch.cuda.stream(st):
                b.copy_(gamma * self.fcur - (self.ycur - self.ahist))
            if m == 0 and call_setup:
                jcur = C.c_int()
                self.H.check(self.H.lib().ampe_precond_setup(self.sim.h, self.t + self.h,
                                                             self._p(self.ycur), gamma,
                                                             C.byref(jcur)))
                assert jcur.value == 1
                self.gamma_setup = gamma
                self.nstlp = self.nst
                self.crate = 1.0
                did_setup = True
            ls = self.spgmr(gamma, b, dl, delta)
            if ls == 2 or (ls == 1 and m > 0):
                return 1, did_setup
            d = self.wrms(dl)
            self.axpby(1.0, self.ycur, 1.0, dl, self.ycur)
            if m > 0:
                self.crate = max(0.3 * self.crate, d / delp)
            if d * min(1.0, self.crate) / s.nlscoef <= 1.0:
                return 0, did_setup
            if m >= 1 and d > 2.0 * delp:
                return 1, did_setup
            delp = d
        return 1, did_setup

    def step(self):
        import torch
        st = self.stream
        s = self.s
        self.H.check(self.H.lib().ampe_sim_reset_ref_concentrations(self.sim.h))
        nrmfac = math.sqrt(self.nc * self.ncell * self.cv)
        delta = s.eplifac * s.nlscoef * nrmfac
        force = False
        while True:
            h = self.h
            with torch.cuda.stream(st):
                if not self.have_prev:
                    self.ypred = self.y.clone()
                    gamma = h
                    self.ahist = self.y.clone()
                else:
                    w = h / self.hprev
                    self.ypred = self.y + w * (self.y - self.yprev)
                    den = 1.0 + 2.0 * w
                    gamma = h * (1.0 + w) / den
                    self.ahist = ((1.0 + w) * (1.0 + w) / den) * self.y - (w * w / den) * self.yprev
            r, did_setup = self.newton(gamma, delta, force)
            assert self.H.lib().ampe_sim_kks_status(self.sim.h, None) == 0
            if r == 0:
                acor, corr = self.zeros(), self.zeros()
                self.axpby(1.0, self.ycur, -1.0, self.ypred, acor)
                self.H.check(self.H.lib().ampe_apply_projection(
                    self.sim.h, self._p(self.ycur), self._p(corr), self._p(acor)))
                self.axpby(1.0, self.ycur, 1.0, corr, self.ycur)
                self.yprev, self.y = self.y, self.ycur
                self.have_prev = True
                self.hprev = h
                self.t += h
                self.nst += 1
                self.h = min(2.0 * h, s.dt)
                self.set_state(self.y)
                return
            self.ncfn += 1
            if not did_setup:
                force = True
                continue
            self.h = 0.25 * h
            force = True
            assert self.h >= s.dt * 1e-8


def _run(dev, case):
    import torch
    from ampe_amd import hip as H
    if case == "two_grain_redblack":
        n = 48
        p = P.two_grain_params(ndim=2)
        p.quad_ceq[0], p.quad_ceq[1] = 0.40, 0.45
        phi, c, q = P.two_grain_state(n=n, radius=n / 4, ndim=2)
        s = P.solver_defaults(maxl=10, max_newton=5)
        shape, nsteps = (n, n), 5
    else:
        n = 16
        p = P.polycrystal_params(seed=1002)
        phi, c, q = P.polycrystal_state(n=n, seed=1002, ngrains=4)
        s = P.solver_defaults(dt=1e-7, maxl=10, max_newton=5, atol=1e-6, rtol=1e-8,
                              precond=P.PRECOND_MG, precision=P.FAC_FP32, mg_coarse_sweeps=16)
        shape, nsteps = (n, n, n), 3
    box = H.boxes_of(shape, len(shape))
    ref = H.Sim(p, s, [box], periodic=box)
    ref.set_state(phi, c, q)
    cb = H.Sim(p, s, [box], periodic=box)
    cb.set_state(phi, c, q)  # the initial phase concentrations
    layout = H.Level(len(shape), [box], 1, periodic=box)
    host = HostBDF(H, cb, p, s, layout, dev)
    comps = [phi] + ([c] if p.with_conc else []) + ([q[m] for m in range(4)] if p.evolve_quat
                                                     else [])
    y0 = np.concatenate([layout.to_ghosted(v) for v in comps])
    assert y0.size == host.nv
    with torch.cuda.stream(host.stream):
        y = torch.from_numpy(y0).to(dev)
    host.set_state(y)
    for step in range(nsteps):
        rc, st = ref.advance(1)
        assert rc == 0
        host.step()
        torch.cuda.synchronize()
        assert (host.nni, host.nli, host.ncfn) == (st["nni"], st["nli"], st["ncfn"]), \
            (step, host.nni, host.nli, host.ncfn, st)
        assert abs(host.t - st["t"]) <= 1e-15 * st["t"]
        fr = ref.get_state(shape)
        yh = host.y.cpu().numpy().reshape(host.nc, -1)
        for k, name in enumerate(["phi"] + (["c"] if p.with_conc else []) +
                                 ([f"q{m}" for m in range(4)] if p.evolve_quat else [])):
            got = layout.ghost_view(yh[k], 0)
            g = layout.ghost
            got = got[g:-g, g:-g, g:-g] if len(shape) == 3 else got[g:-g, g:-g]
            if name == "phi":
                want = fr[0]
            elif name == "c":
                want = fr[1]
            else:
                want = fr[2][int(name[1])]
            e = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert e < 1e-10, (step, name, e)


@pytest.mark.parametrize("case", ["two_grain_redblack", "calphad_mg_fp32"])
def test_host_loop_over_callbacks_reproduces_sim(dev, case):
    _run(dev, case)
