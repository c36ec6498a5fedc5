// BDF Newton-Krylov integrator on one patch level, device resident.
//
// The host loop is the QuatIntegrator / CVODE callback structure
// (source/QuatIntegrator.cc:2096-2233, 3134-3771; CVODE options
// source/samrai/CVODESolver.cc:150-247): RHS evaluations with lagged
// orientation face gradients and composition coefficients in Jv, SPGMR with
// modified Gram-Schmidt and left block preconditioning, difference-quotient
// Jv, Newton with CVODE's convergence-rate test, projection of q onto |q| = 1
// after each step.  Step size is the imposed dt (BDF1 first, variable-step
// BDF2 after); a Newton failure with a fresh preconditioner cuts h by 4 and h
// regrows by 2 per accepted step.  All vectors live in HBM; only Hessenberg
// scalars cross to the host.
#include "sim.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace ampe {

static int dalloc(double** p, long long n)
{
   if (n <= 0) n = 1;
   hipError_t e = hipMalloc(p, sizeof(double) * (size_t)n);
   if (e != hipSuccess) {
      set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
      return -1;
   }
   e = hipMemset(*p, 0, sizeof(double) * (size_t)n);
   return e == hipSuccess ? 0 : -1;
}

static int falloc(float** p, long long n)
{
   if (n <= 0) n = 1;
   hipError_t e = hipMalloc(p, sizeof(float) * (size_t)n);
   if (e != hipSuccess) {
      set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
      return -1;
   }
   e = hipMemset(*p, 0, sizeof(float) * (size_t)n);
   return e == hipSuccess ? 0 : -1;
}

// First step: dt_init when 0 < dt_init < dt, else dt.
static double first_step(const ampe_solver_params& s)
{
   return (s.dt_init > 0.0 && s.dt_init < s.dt) ? s.dt_init : s.dt;
}

static int sim_finish(Sim* S, Sim** out);

static int sim_check(const ampe_model_params& p, const ampe_solver_params& s)
{
   if (s.adaptive && (s.max_order < 1 || s.max_order > 2)) {
      set_error("sim_create: the adaptive BDF needs max_order 1 or 2");
      return -1;
   }
   if (s.maxl < 1 || s.maxl > 30) {
      set_error("sim_create: maxl must be in [1, 30]");
      return -1;
   }
   if (s.gstype != AMPE_GS_MODIFIED && (s.gstype != AMPE_GS_CLASSICAL || s.maxl > kMaxCgs - 2)) {
      set_error("sim_create: gstype must be modified, or classical with maxl <= 14");
      return -1;
   }
   if (p.evolve_quat && p.qlen != 4) {
      set_error("sim_create: qlen must be 4");
      return -1;
   }
   if (p.phi_mobility_type == AMPE_PHI_MOB_KIM && !p.with_conc) {
      set_error("sim_create: Kim's phase mobility needs a composition field");
      return -1;
   }
   if (p.conc_rhs_form != AMPE_CONC_RHS_KKS && p.conc_rhs_form != AMPE_CONC_RHS_EBS) {
      set_error("sim_create: unknown composition rhs form");
      return -1;
   }
   if (p.conc_rhs_form == AMPE_CONC_RHS_EBS && p.ebs_stencil != AMPE_EBS_REGULAR &&
       p.ebs_stencil != AMPE_EBS_ISOTROPIC && p.ebs_stencil != AMPE_EBS_ORDER4) {
      set_error("sim_create: unknown EBS stencil");
      return -1;
   }
   if (p.conc_diffusion_type != AMPE_CONC_DIFF_TEMPERATURE &&
       (p.conc_diffusion_type != AMPE_CONC_DIFF_COMPOSITION ||
        p.conc_rhs_form != AMPE_CONC_RHS_EBS || p.ebs_stencil == AMPE_EBS_ISOTROPIC ||
        p.fe_model == AMPE_FE_QUADRATIC_TERNARY_3PH)) {
      set_error("sim_create: composition-dependent diffusion needs the two-phase EBS form "
                "with the regular or order4 stencil");
      return -1;
   }
   const bool tern = p.fe_model == AMPE_FE_QUADRATIC_TERNARY_3PH;
   if (tern && (p.norderp < 3 || !p.with_conc || p.evolve_quat ||
                p.conc_rhs_form != AMPE_CONC_RHS_EBS || p.ebs_stencil != AMPE_EBS_REGULAR ||
                p.norderp_A < 1 || p.norderp_A > p.norderp - 1)) {
      set_error("sim_create: the three-phase ternary model needs norderp >= 3, 1 <= norderp_A "
                "<= norderp - 1, a composition field with the regular EBS form and no "
                "orientation field");
      return -1;
   }
   if (p.norderp > 1 && !tern && (p.with_conc || p.evolve_quat)) {
      set_error("sim_create: several order parameters need the three-phase ternary model or "
                "no composition and orientation fields");
      return -1;
   }
   if (p.norderp > 1 && s.precond != AMPE_PRECOND_REDBLACK) {
      set_error("sim_create: several order parameters run the red-black preconditioner");
      return -1;
   }
   return 0;
}

int sim_create(const ampe_model_params& p, const ampe_solver_params& s, const ampe_box* boxes,
               int npatch, const ampe_box* periodic, Sim** out)
{
   // one level batches the preconditioner's blocks, at most eight
   if (p.norderp + 2 > 8) {
      set_error("sim_create: several order parameters on one level: norderp <= 6 (the "
                "hierarchy integrator has no such limit)");
      return -1;
   }
   if (sim_check(p, s)) return -1;
   Sim* S = new Sim();
   S->p = p;
   S->s = s;
   // two ghost cells for the fourth-order composition stencil, one otherwise
   // (QuatModelParameters::nghosts_required, source/QuatModelParameters.h:305-311)
   const int ghost =
       (p.with_conc && p.conc_rhs_form == AMPE_CONC_RHS_EBS && p.ebs_stencil == AMPE_EBS_ORDER4)
           ? 2
           : 1;
   if (level_create(p.ndim, boxes, npatch, ghost, periodic, &S->L)) {
      delete S;
      return -1;
   }
   return sim_finish(S, out);
}

// Vectors, work arrays, the preconditioner's hierarchy and the stream of a
// sim whose level S->L is set.
static int sim_finish(Sim* S, Sim** out)
{
   const ampe_model_params& p = S->p;
   const ampe_solver_params& s = S->s;
   const Level& L = *S->L;
   S->nord = p.norderp > 1 ? p.norderp : 1;
   S->tern = p.fe_model == AMPE_FE_QUADRATIC_TERNARY_3PH;
   S->ncs = p.with_conc ? (S->tern ? 2 : 1) : 0;
   S->nc = S->nord + S->ncs + (p.evolve_quat ? p.qlen : 0);
   S->ic = p.with_conc ? S->nord : -1;
   S->iq = p.evolve_quat ? S->nord + S->ncs : -1;
   S->ng = L.ncg;
   S->nv = (long long)S->nc * S->ng;
   S->nf = L.nside_tot;
   int ql = p.evolve_quat ? p.qlen : 1;
   int rc = 0;
   double** vecs[] = {&S->y, &S->yprev, &S->ypred, &S->ycur, &S->ahist, &S->fcur, &S->b,
                      &S->del, &S->ewt, &S->work, &S->work2, &S->jv, &S->xk, &S->corr,
                      &S->acor};
   for (auto v : vecs) rc |= dalloc(v, S->nv);
   if (s.adaptive) {
      rc |= dalloc(&S->z1, S->nv);
      rc |= dalloc(&S->z2, S->nv);
   }
   S->V.assign(s.maxl + 1, nullptr);
   for (auto& v : S->V) rc |= dalloc(&v, S->nv);
   double** cells[] = {&S->cl, &S->cs, &S->clr, &S->csr, &S->Cphi, &S->sqrtm, &S->Mphi,
                       &S->phid};
   for (auto v : cells) rc |= dalloc(v, S->ng);
   rc |= dalloc(&S->fac_u, S->ng * ql);
   rc |= dalloc(&S->fac_e, S->nv);
   rc |= dalloc(&S->fac_r, S->nv);
   double** sides[] = {&S->n2, &S->Dc, &S->Dphi, &S->Dconc, &S->fc, &S->fcq};
   for (auto v : sides) rc |= dalloc(v, S->nf);
   if (S->tern) {
      rc |= dalloc(&S->ph6, 6 * S->ng);
      rc |= dalloc(&S->Db, S->nf);
   }
   if (s.precision == AMPE_FAC_FP32) {
      rc |= falloc(&S->f_u, S->nv);
      rc |= falloc(&S->f_e, S->nv);
      rc |= falloc(&S->f_r, S->nv);
      rc |= falloc(&S->f_f, S->nv);
      rc |= falloc(&S->f_Cphi, S->ng);
      rc |= falloc(&S->f_Mphi, S->ng);
      rc |= falloc(&S->f_sqrtm, S->ng);
      rc |= falloc(&S->f_Dconc, S->nf);
      rc |= falloc(&S->f_fc, S->nf);
   }
   if (!rc && s.precond == AMPE_PRECOND_MG) {
      int maxlev = s.mg_max_levels > 0 ? s.mg_max_levels : 32;
      if (s.precision == AMPE_FAC_FP32) {
         rc = mg_create<float>(L, p.dx, ql, maxlev, 2, &S->mgf);
         if (!rc && p.evolve_quat && !s.mg_quat_natural) {
            rc = mg_enable_qsplit<float>(*S->mgf, nullptr);
            if (!rc && hipDeviceSynchronize() != hipSuccess) rc = -1;
         }
      } else
         rc = mg_create<double>(L, p.dx, ql, maxlev, 2, &S->mgd);
   } else if (s.precond != AMPE_PRECOND_REDBLACK) {
      set_error("sim_create: unknown preconditioner");
      rc = -1;
   }
   if (!rc && hipMalloc(&S->kks_acc, sizeof(long long) * kKksAcc) != hipSuccess) rc = -1;
   if (!rc && hipMemset(S->kks_acc, 0, sizeof(long long) * kKksAcc) != hipSuccess) rc = -1;
   if (rc || hipStreamCreate(&S->stream) != hipSuccess ||
       hipEventCreate(&S->ev0) != hipSuccess || hipEventCreate(&S->ev1) != hipSuccess) {
      if (!rc) set_error("sim_create: stream/event creation failed");
      sim_destroy(S);
      return -1;
   }
   S->h = first_step(s);
   *out = S;
   return 0;
}

int sim_create_hier(const ampe_model_params& p, const ampe_solver_params& s, Hier* H, Sim** out)
{
   if (sim_check(p, s)) return -1;
   if (!H) {
      set_error("sim_create_hier: null hierarchy");
      return -1;
   }
   if (s.precond != AMPE_PRECOND_REDBLACK || s.graph_psolve) {
      set_error("sim_create_hier: the hierarchy runs the FAC preconditioner over levels "
                "(precond redblack, FP64 or FP32, no graph capture)");
      return -1;
   }
   const bool tern = p.fe_model == AMPE_FE_QUADRATIC_TERNARY_3PH;
   if (p.with_conc && !tern && p.conc_rhs_form != AMPE_CONC_RHS_KKS) {
      set_error("sim_create_hier: the hierarchy supports the KKS composition form and the "
                "three-phase ternary EBS form");
      return -1;
   }
   if (s.kks_guess == AMPE_GUESS_SURROGATE) {
      set_error("sim_create_hier: the previous-step KKS guess only");
      return -1;
   }
   Sim* S = new Sim();
   S->p = p;
   S->s = s;
   S->hier = H;
   S->L = hier_composite(H);
   double sw = 0.0;
   if (hier_weight_sum(H, &sw)) {
      delete S;
      return -1;
   }
   const int nord = p.norderp > 1 ? p.norderp : 1;
   const int ncs = p.with_conc ? (tern ? 2 : 1) : 0;
   S->wsum = sw * (double)(nord + ncs + (p.evolve_quat ? p.qlen : 0));
   return sim_finish(S, out);
}

void sim_destroy(Sim* S)
{
   if (!S) return;
   double* all[] = {S->y, S->yprev, S->ypred, S->ycur, S->ahist, S->fcur, S->b, S->del,
                    S->ewt, S->work, S->work2, S->jv, S->xk, S->corr, S->acor, S->cl, S->cs,
                    S->clr, S->csr, S->Cphi, S->sqrtm, S->Mphi, S->fac_u, S->fac_e, S->fac_r,
                    S->n2,
                    S->Dc, S->Dphi, S->Dconc, S->fc, S->fcq, S->phid, S->z1, S->z2,
                    S->ph6, S->Db};
   for (double* v : all) (void)hipFree(v);
   for (double* v : S->V) (void)hipFree(v);
   float* fall[] = {S->f_u,    S->f_e,     S->f_r,     S->f_f,   S->f_Cphi,
                    S->f_sqrtm, S->f_Dconc, S->f_fc, S->f_Mphi};
   for (float* v : fall) (void)hipFree(v);
   (void)hipFree(S->kks_acc);
   (void)hipFree(S->guess.list);
   (void)hipFree(S->guess.count);
   mg_destroy(S->mgf);
   mg_destroy(S->mgd);
   if (S->pgraph) (void)hipGraphExecDestroy(S->pgraph);
   if (S->stream) (void)hipStreamDestroy(S->stream);
   if (S->ev0) (void)hipEventDestroy(S->ev0);
   if (S->ev1) (void)hipEventDestroy(S->ev1);
   if (!S->hier) level_destroy(S->L);
   delete S;
}

// ------------------------------------------------------------ host I/O
static int scatter_interior(Sim* S, const double* host, double* dev)
{
   const Level& L = *S->L;
   std::vector<double> buf((size_t)L.ncg, 0.0);
   AMPE_HIP_CHECK(hipMemcpy(buf.data(), dev, sizeof(double) * L.ncg, hipMemcpyDeviceToHost));
   for (const PatchDev& d : L.hp) {
      long long n = d.int_off;
      for (int k = 0; k < d.nz; k++)
         for (int j = 0; j < d.ny; j++)
            for (int i = 0; i < d.nx; i++)
               buf[(size_t)(d.cell_off + (i + L.g) + (long long)(j + L.g) * d.sy +
                            (long long)(k + L.gz) * d.sz)] = host[n++];
   }
   AMPE_HIP_CHECK(hipMemcpy(dev, buf.data(), sizeof(double) * L.ncg, hipMemcpyHostToDevice));
   return 0;
}

static int gather_interior(Sim* S, const double* dev, double* host)
{
   const Level& L = *S->L;
   std::vector<double> buf((size_t)L.ncg);
   AMPE_HIP_CHECK(hipMemcpy(buf.data(), dev, sizeof(double) * L.ncg, hipMemcpyDeviceToHost));
   for (const PatchDev& d : L.hp) {
      long long n = d.int_off;
      for (int k = 0; k < d.nz; k++)
         for (int j = 0; j < d.ny; j++)
            for (int i = 0; i < d.nx; i++)
               host[n++] = buf[(size_t)(d.cell_off + (i + L.g) + (long long)(j + L.g) * d.sy +
                                        (long long)(k + L.gz) * d.sz)];
   }
   return 0;
}

static int kks(Sim* S, const double* y)
{
   const double* c = y + (long long)S->ic * S->ng;
   if (S->tern) return launch_kks_tern3(*S->L, S->p, S->nord, y, c, S->ph6, S->stream);
   KksGuess g;
   if (S->s.kks_guess == AMPE_GUESS_SURROGATE) {
      if (!S->guess.surrogate) {
         set_error("kks_guess is the surrogate but no surrogate was set");
         return -1;
      }
      g = S->guess;
   }
   return launch_kks(*S->L, S->p, y, c, S->clr, S->csr, S->cl, S->cs, g, S->kks_acc, S->stream);
}

// Read the KKS device counters into the stats; -1 if a cell failed.
static int kks_collect(Sim* S)
{
   if (S->ic < 0) return 0;
   AMPE_HIP_CHECK(hipMemcpyAsync(S->L->hpinned, S->kks_acc, sizeof(long long) * kKksAcc,
                                 hipMemcpyDeviceToHost, S->stream));
   AMPE_HIP_CHECK(hipStreamSynchronize(S->stream));
   long long h[kKksAcc];
   std::memcpy(h, S->L->hpinned, sizeof(h));
   S->st.kks_iters = h[0];
   S->st.kks_calls = h[3];
   S->st.kks_fallbacks = h[4];
   S->st.kks_surrogate_starts = h[5];
   S->st.kks_cells = S->st.kks_calls * S->L->ncg;
   if (h[2]) {
      set_error("KKS Newton did not converge in a cell");
      return -1;
   }
   return 0;
}

int sim_set_state(Sim* S, const double* phi, const double* c, const double* q)
{
   const Level& L = *S->L;
   hipStream_t st = S->stream;
   AMPE_HIP_CHECK(hipMemset(S->y, 0, sizeof(double) * S->nv));
   for (int m = 0; m < S->nord; m++)
      if (scatter_interior(S, phi + (long long)m * L.ncell, S->y + (long long)m * S->ng))
         return -1;
   if (S->ic >= 0) {
      if (!c) {
         set_error("set_state: model has a composition field");
         return -1;
      }
      for (int m = 0; m < S->ncs; m++)
         if (scatter_interior(S, c + (long long)m * L.ncell,
                              S->y + (long long)(S->ic + m) * S->ng))
            return -1;
   }
   if (S->iq >= 0) {
      if (!q) {
         set_error("set_state: model has an orientation field");
         return -1;
      }
      for (int m = 0; m < S->p.qlen; m++)
         if (scatter_interior(S, q + (long long)m * L.ncell,
                              S->y + (long long)(S->iq + m) * S->ng))
            return -1;
   }
   if (ghost_fill(L, S->y, S->nc, st)) return -1;
   std::memset(&S->st, 0, sizeof(S->st));
   if (S->ic >= 0) {
      const double* cc = S->y + (long long)S->ic * S->ng;
      AMPE_HIP_CHECK(hipMemcpyAsync(S->clr, cc, sizeof(double) * S->ng,
                                    hipMemcpyDeviceToDevice, st));
      AMPE_HIP_CHECK(hipMemcpyAsync(S->csr, cc, sizeof(double) * S->ng,
                                    hipMemcpyDeviceToDevice, st));
      AMPE_HIP_CHECK(hipMemsetAsync(S->kks_acc, 0, sizeof(long long) * kKksAcc, st));
      if (kks(S, S->y)) return -1;
      if (kks_collect(S)) return -1;
      std::memset(&S->st, 0, sizeof(S->st));
      AMPE_HIP_CHECK(hipMemsetAsync(S->kks_acc, 0, sizeof(long long) * kKksAcc, st));
   }
   S->t = 0.0;
   S->nst = 0;
   S->nstlp = 0;
   S->have_prev = false;
   S->h = first_step(S->s);
   S->bdf = BdfCtl{};
   if (vec_ewt(S->nv, S->y, S->s.rtol, S->s.atol, S->ewt, st)) return -1;
   AMPE_HIP_CHECK(hipStreamSynchronize(st));
   return 0;
}

int sim_set_state_composite(Sim* S, const double* y_host)
{
   if (!S->hier) {
      set_error("set_state_composite: not a hierarchy sim");
      return -1;
   }
   hipStream_t st = S->stream;
   AMPE_HIP_CHECK(hipMemcpy(S->y, y_host, sizeof(double) * S->nv, hipMemcpyHostToDevice));
   int ql = S->iq >= 0 ? S->p.qlen : 0;
   if (hier_sim_sync_fill(S->hier, S->y, S->nc, S->nord, S->ic, S->ncs, S->iq, ql, st)) return -1;
   std::memset(&S->st, 0, sizeof(S->st));
   if (S->ic >= 0) {
      const double* cc = S->y + (long long)S->ic * S->ng;
      AMPE_HIP_CHECK(hipMemcpyAsync(S->clr, cc, sizeof(double) * S->ng,
                                    hipMemcpyDeviceToDevice, st));
      AMPE_HIP_CHECK(hipMemcpyAsync(S->csr, cc, sizeof(double) * S->ng,
                                    hipMemcpyDeviceToDevice, st));
      AMPE_HIP_CHECK(hipMemsetAsync(S->kks_acc, 0, sizeof(long long) * kKksAcc, st));
      if (kks(S, S->y)) return -1;
      if (kks_collect(S)) return -1;
      std::memset(&S->st, 0, sizeof(S->st));
      AMPE_HIP_CHECK(hipMemsetAsync(S->kks_acc, 0, sizeof(long long) * kKksAcc, st));
   }
   S->t = 0.0;
   S->nst = 0;
   S->nstlp = 0;
   S->have_prev = false;
   S->h = first_step(S->s);
   S->bdf = BdfCtl{};
   if (vec_ewt(S->nv, S->y, S->s.rtol, S->s.atol, S->ewt, st)) return -1;
   AMPE_HIP_CHECK(hipStreamSynchronize(st));
   return 0;
}

// The integrator restarted at time t after a set_state (a regrid:
// QuatIntegrator::resetIntegrator, source/QuatIntegrator.cc:1676-1730, with
// setSundialsOptions' initial time and step, :1571-1580); h0 > 0 is the
// first step, else the solver's own first step.
int sim_set_time(Sim* S, double t, double h0)
{
   if (!(t >= 0.0) || h0 != h0) {
      set_error("sim_set_time: t must be >= 0");
      return -1;
   }
   S->t = t;
   if (h0 > 0.0) S->h = h0;
   return 0;
}

// The state with every finer level coarsened onto the cells it covers (the
// output QuatIntegrator coarsens after each step, source/QuatIntegrator.cc:
// 2190-2196); the integrator's own state is left as it is.
int sim_get_state_composite(Sim* S, double* y_host)
{
   if (!S->hier) {
      set_error("get_state_composite: not a hierarchy sim");
      return -1;
   }
   hipStream_t st = S->stream;
   AMPE_HIP_CHECK(hipMemcpyAsync(S->work, S->y, sizeof(double) * S->nv, hipMemcpyDeviceToDevice,
                                 st));
   if (hier_coarsen(*S->hier, S->work, S->nord, 0, st)) return -1;
   if (S->ic >= 0 &&
       hier_coarsen(*S->hier, S->work + (long long)S->ic * S->ng, S->ncs, 0, st))
      return -1;
   if (S->iq >= 0 &&
       hier_coarsen(*S->hier, S->work + (long long)S->iq * S->ng, S->p.qlen, 1, st))
      return -1;
   AMPE_HIP_CHECK(hipMemcpyAsync(y_host, S->work, sizeof(double) * S->nv, hipMemcpyDeviceToHost,
                                 st));
   AMPE_HIP_CHECK(hipStreamSynchronize(st));
   return 0;
}

int sim_get_state(Sim* S, double* phi, double* c, double* q, double* cl, double* cs)
{
   const Level& L = *S->L;
   AMPE_HIP_CHECK(hipStreamSynchronize(S->stream));
   for (int m = 0; phi && m < S->nord; m++)
      if (gather_interior(S, S->y + (long long)m * S->ng, phi + (long long)m * L.ncell)) return -1;
   for (int m = 0; c && m < S->ncs; m++)
      if (gather_interior(S, S->y + (long long)(S->ic + m) * S->ng, c + (long long)m * L.ncell))
         return -1;
   if (q && S->iq >= 0)
      for (int m = 0; m < S->p.qlen; m++)
         if (gather_interior(S, S->y + (long long)(S->iq + m) * S->ng, q + (long long)m * L.ncell))
            return -1;
   // the three-phase model: cl = liquid, cs = solid A, two species each
   for (int m = 0; S->ic >= 0 && m < S->ncs; m++) {
      const double* a = S->tern ? S->ph6 + (long long)m * S->ng : S->cl;
      const double* b = S->tern ? S->ph6 + (long long)(2 + m) * S->ng : S->cs;
      if (cl && gather_interior(S, a, cl + (long long)m * L.ncell)) return -1;
      if (cs && gather_interior(S, b, cs + (long long)m * L.ncell)) return -1;
   }
   return 0;
}


// ------------------------------------------------------------ callbacks
// evaluateRHSFunction (source/QuatIntegrator.cc:3134-3298)
// ------------------------------------------------------------ hierarchy
// The callbacks over all levels of a hierarchy: setCoefficients coarsens the
// iterate and fillScratch fills every level's ghosts (LINEAR_REFINE on the
// coarse-fine boundary), then the composite right-hand sides
// (source/QuatIntegrator.cc:2994-3083, 2873-2957, 3134-3298); the
// preconditioner blocks are FAC cycles over the levels.
static int hier_rhs_eval(Sim* S, double* y, double* ydot, bool fd)
{
   const ampe_model_params& p = S->p;
   hipStream_t st = S->stream;
   bool recompute = !fd || !S->s.lag_quat_sidegrad;
   S->st.nfe++;
   int ql = S->iq >= 0 ? p.qlen : 0;
   if (hier_sim_sync_fill(S->hier, y, S->nc, S->nord, S->ic, S->ncs, S->iq, ql, st)) return -1;
   const double* phi = y;
   const double* q = S->iq >= 0 ? y + (long long)S->iq * S->ng : nullptr;
   const double* c = S->ic >= 0 ? y + (long long)S->ic * S->ng : nullptr;
   if (S->tern) {
      if (kks(S, y)) return -1;
      if (hier_phase_rhs_tern3(*S->hier, p, S->nord, phi, S->ph6, ydot, st)) return -1;
      if (recompute && hier_ebs3_coeffs(*S->hier, p, S->nord, phi, S->Dc, S->Dphi, S->Db, st))
         return -1;
      return hier_conc_rhs_ebs3(*S->hier, p, S->ph6, S->Dc, S->Dphi, S->Db,
                                ydot + (long long)S->ic * S->ng, st);
   }
   if (S->nord > 1) return hier_phase_rhs_multiorder(*S->hier, p, S->nord, phi, ydot, st);
   if (q && recompute && hier_quat_face_norm2(*S->hier, p, q, S->n2, st)) return -1;
   if (c && kks(S, y)) return -1;
   if (hier_phase_rhs(*S->hier, p, phi, q, S->cl, S->cs, ydot, st)) return -1;
   if (q && hier_quat_rhs(*S->hier, p, phi, q, S->n2, S->fcq, ydot + (long long)S->iq * S->ng,
                          st))
      return -1;
   if (c) {
      if (recompute && hier_conc_coeffs(*S->hier, p, phi, S->cl, S->cs, S->Dc, S->Dphi, st))
         return -1;
      if (hier_conc_rhs(*S->hier, p, c, phi, S->Dc, S->Dphi, ydot + (long long)S->ic * S->ng,
                        st))
         return -1;
   }
   return 0;
}

static int hier_psetup(Sim* S, double* y, double gamma)
{
   const ampe_model_params& p = S->p;
   hipStream_t st = S->stream;
   S->st.npe++;
   int ql = S->iq >= 0 ? p.qlen : 0;
   if (hier_sim_sync_fill(S->hier, y, S->nc, S->nord, S->ic, S->ncs, S->iq, ql, st)) return -1;
   const double* phi = y;
   const double* q = S->iq >= 0 ? y + (long long)S->iq * S->ng : nullptr;
   const double* c = S->ic >= 0 ? y + (long long)S->ic * S->ng : nullptr;
   if (q && hier_quat_face_norm2(*S->hier, p, q, S->n2, st)) return -1;
   if (c && kks(S, y)) return -1;
   // several orders: the well scale passed is 0, so C = 1
   ampe_model_params pw = p;
   if (S->nord > 1) pw.phi_well_scale = 0.0;
   if (hier_precond_coeffs(*S->hier, pw, gamma, phi, S->cl, S->cs, S->n2, S->Mphi, S->Cphi,
                           q ? S->fc : nullptr, q ? S->sqrtm : nullptr, st))
      return -1;
   if (c && S->tern) {
      // D_L + D_A + D_B of the coarsened face diffusivities
      if (vec_axpby(S->nf, 1.0, S->Dc, 1.0, S->Dphi, S->Dconc, st)) return -1;
      if (vec_axpby(S->nf, 1.0, S->Dconc, 1.0, S->Db, S->Dconc, st)) return -1;
      if (vec_axpby(S->nf, -gamma, S->Dconc, 0.0, S->Dconc, S->Dconc, st)) return -1;
   } else if (c && vec_axpby(S->nf, -gamma, S->Dc, 0.0, S->Dc, S->Dconc, st)) {
      return -1;
   }
   if (S->s.precision == AMPE_FAC_FP32) {
      if (vec_to_f32(S->ng, S->Cphi, S->f_Cphi, st)) return -1;
      if (p.phi_mobility_type == AMPE_PHI_MOB_KIM && vec_to_f32(S->ng, S->Mphi, S->f_Mphi, st))
         return -1;
      if (c && vec_to_f32(S->nf, S->Dconc, S->f_Dconc, st)) return -1;
      if (q) {
         if (vec_to_f32(S->nf, S->fc, S->f_fc, st)) return -1;
         if (vec_to_f32(S->ng, S->sqrtm, S->f_sqrtm, st)) return -1;
      }
   }
   S->gamma_setup = gamma;
   return 0;
}

static int hier_psolve(Sim* S, const double* r, double* z, double delta)
{
   const ampe_model_params& p = S->p;
   hipStream_t st = S->stream;
   long long ng = S->ng;
   S->st.nps++;
   int ql = S->iq >= 0 ? p.qlen : 0;
   int iq = S->iq >= 0 ? S->iq : S->nc;
   int ret = 0;
   if (!r || !z) {
      set_error("hier_psolve: needs the input and output vectors");
      return -1;
   }
   ampe_fac_params fp{S->s.fac_max_cycles, 1, 1, S->s.fac_sweeps, S->s.fac_tol};
   if (S->s.precision == AMPE_FAC_FP32) {
      // the same blocks in FP32: f = r (orientation blocks / sqrt(M), FP64
      // division), u = 0, FAC in FP32, z = u (orientation blocks * sqrt(M))
      if (psolve_in<float>(ng, S->nc, iq, ql, r, S->sqrtm, S->f_f, nullptr, S->f_u, st))
         return -1;
      const bool kim = p.phi_mobility_type == AMPE_PHI_MOB_KIM;
      HBlock<float> ph{0,
                       nullptr,
                       (float)(-S->gamma_setup * p.epsilon_phi * p.epsilon_phi),
                       S->f_Cphi,
                       1.0f,
                       kim ? S->f_Mphi : nullptr,
                       (float)p.phi_mobility,
                       nullptr,
                       0.0f,
                       1};
      for (int m = 0; m < S->nord; m++) {
         const long long off = (long long)m * ng;
         int r0 = hier_fac_solve_f32(*S->hier, ph, fp, S->f_f + off, S->f_u + off, 1, delta,
                                     nullptr, nullptr, st);
         if (r0 < 0) return -1;
         ret |= r0;
      }
      if (S->iq >= 0) {
         ampe_fac_params qp{S->s.quat_max_cycles, 1, 1, S->s.quat_sweeps, -1.0};
         HBlock<float> qb{1,       S->f_fc, 0.0f,        nullptr, 1.0f, nullptr, 1.0f,
                          S->f_sqrtm, (float)S->gamma_setup, 0};
         const long long off = (long long)S->iq * ng;
         int r1 = hier_fac_solve_f32(*S->hier, qb, qp, S->f_f + off, S->f_u + off, ql, delta,
                                     nullptr, nullptr, st);
         if (r1 < 0) return -1;
         ret |= r1;
      }
      for (int m = 0; m < S->ncs; m++) {
         HBlock<float> co{0, S->f_Dconc, 0.0f, nullptr, 1.0f, nullptr, (float)p.conc_mobility,
                          nullptr, 0.0f, 1};
         const long long off = (long long)(S->ic + m) * ng;
         int r2 = hier_fac_solve_f32(*S->hier, co, fp, S->f_f + off, S->f_u + off, 1, delta,
                                     nullptr, nullptr, st);
         if (r2 < 0) return -1;
         ret |= r2;
      }
      if (psolve_out<float>(ng, S->nc, iq, ql, S->f_u, S->sqrtm, nullptr, z, st)) return -1;
      return ret;
   }
   // z = 0, fac_u = orientation blocks of r / sqrt(M)
   if (psolve_in<double>(ng, S->nc, iq, ql, r, S->sqrtm, nullptr, S->fac_u, z, st)) return -1;
   ampe_hier_block ph{0,       nullptr, -S->gamma_setup * p.epsilon_phi * p.epsilon_phi,
                      S->Cphi, 1.0,
                      p.phi_mobility_type == AMPE_PHI_MOB_KIM ? S->Mphi : nullptr,
                      p.phi_mobility, nullptr, 0.0, 1};
   // one phase block per order (PhasePrecondSolve, source/QuatIntegrator.cc:3427-3444)
   for (int m = 0; m < S->nord; m++) {
      int r0 = hier_fac_solve(*S->hier, ph, fp, r + (long long)m * ng, z + (long long)m * ng, 1,
                              delta, nullptr, nullptr, st);
      if (r0 < 0) return -1;
      ret |= r0;
   }
   if (S->iq >= 0) {
      ampe_fac_params qp{S->s.quat_max_cycles, 1, 1, S->s.quat_sweeps, -1.0};
      ampe_hier_block qb{1, S->fc, 0.0, nullptr, 1.0, nullptr, 1.0, S->sqrtm, S->gamma_setup, 0};
      int r1 = hier_fac_solve(*S->hier, qb, qp, S->fac_u, z + (long long)S->iq * ng, ql, delta,
                              nullptr, nullptr, st);
      if (r1 < 0) return -1;
      ret |= r1;
      if (vec_mul_comp(S, z + (long long)S->iq * ng, S->sqrtm, ql)) return -1;
   }
   // one composition block per species
   for (int m = 0; m < S->ncs; m++) {
      ampe_hier_block co{0, S->Dconc, 0.0, nullptr, 1.0, nullptr, p.conc_mobility, nullptr, 0.0,
                         1};
      const long long off = (long long)(S->ic + m) * ng;
      int r2 = hier_fac_solve(*S->hier, co, fp, r + off, z + off, 1, delta, nullptr, nullptr, st);
      if (r2 < 0) return -1;
      ret |= r2;
   }
   return ret;
}

static int rhs_eval(Sim* S, double* y, double* ydot, bool fd)
{
   if (S->hier) return hier_rhs_eval(S, y, ydot, fd);
   const Level& L = *S->L;
   const ampe_model_params& p = S->p;
   hipStream_t st = S->stream;
   bool recompute = !fd || !S->s.lag_quat_sidegrad;
   S->st.nfe++;
   if (ghost_fill(L, y, S->nc, st)) return -1;
   const double* phi = y;
   const double* q = S->iq >= 0 ? y + (long long)S->iq * S->ng : nullptr;
   const double* c = S->ic >= 0 ? y + (long long)S->ic * S->ng : nullptr;
   if (S->tern) {
      // three phases, two species: closed-form KKS, the multi-order phase RHS
      // with the driving force, the EBS composition RHS of the three phases
      if (kks(S, y)) return -1;
      if (launch_phase_rhs_tern3(L, p, S->nord, phi, S->ph6, nullptr, ydot, st)) return -1;
      if (recompute && launch_ebs3_coeffs(L, p, S->nord, phi, S->Dc, S->Dphi, S->Db, st))
         return -1;
      return launch_conc_rhs_ebs3(L, p, S->ph6, S->Dc, S->Dphi, S->Db,
                                  ydot + (long long)S->ic * S->ng, st);
   }
   if (S->nord > 1) return launch_phase_rhs_multiorder(L, p, S->nord, phi, nullptr, ydot, st);
   if (q && recompute && launch_quat_grad(L, p, q, nullptr, S->n2, st)) return -1;
   if (c && kks(S, y)) return -1;
   if (launch_phase_rhs(L, p, phi, q, S->cl, S->cs, ydot, st)) return -1;
   // face coefficients once per face, then the orientation RHS from them
   if (q && launch_quat_face_coef(L, p, phi, S->n2, S->fcq, st)) return -1;
   if (q && launch_quat_rhs_fc(L, p, phi, q, S->fcq, ydot + (long long)S->iq * S->ng, st))
      return -1;
   if (c && p.conc_rhs_form == AMPE_CONC_RHS_EBS) {
      // Dc, Dphi hold the liquid and solid face diffusivities
      if (recompute) {
         if (p.conc_diffusion_type == AMPE_CONC_DIFF_COMPOSITION) {
            if (launch_ebs_coeffs_mob(L, p, phi, S->cl, S->cs, S->Dc, S->Dphi, st)) return -1;
         } else if (launch_ebs_coeffs(L, p, phi, S->Dc, S->Dphi, st)) {
            return -1;
         }
      }
      if (p.ebs_stencil == AMPE_EBS_ISOTROPIC) {
         // the diffusivities follow the lagging of the face data: the
         // isotropic stencil evaluates them from the phase field of the last
         // full evaluation, kept in phid (ghosts filled)
         if (recompute)
            AMPE_HIP_CHECK(hipMemcpyAsync(S->phid, phi, sizeof(double) * S->ng,
                                          hipMemcpyDeviceToDevice, st));
         if (launch_conc_rhs_ebs_iso(L, p, S->phid, S->cl, S->cs,
                                     ydot + (long long)S->ic * S->ng, st))
            return -1;
      } else if (p.ebs_stencil == AMPE_EBS_ORDER4) {
         if (launch_conc_rhs_ebs_4th(L, p, S->cl, S->cs, S->Dc, S->Dphi,
                                     ydot + (long long)S->ic * S->ng, st))
            return -1;
      } else if (launch_conc_rhs_ebs(L, p, S->cl, S->cs, S->Dc, S->Dphi,
                                     ydot + (long long)S->ic * S->ng, st)) {
         return -1;
      }
   } else if (c) {
      if (recompute && launch_conc_coeffs(L, p, phi, S->cl, S->cs, S->Dc, S->Dphi, st)) return -1;
      if (launch_conc_rhs(L, p, c, phi, S->Dc, S->Dphi, ydot + (long long)S->ic * S->ng, st))
         return -1;
   }
   return 0;
}

// CVSpgmrPrecondSet (source/QuatIntegrator.cc:3300-3376)
static int psetup(Sim* S, double* y, double gamma)
{
   if (S->hier) return hier_psetup(S, y, gamma);
   const Level& L = *S->L;
   const ampe_model_params& p = S->p;
   hipStream_t st = S->stream;
   S->st.npe++;
   if (ghost_fill(L, y, S->nc, st)) return -1;
   const double* phi = y;
   const double* q = S->iq >= 0 ? y + (long long)S->iq * S->ng : nullptr;
   const double* c = S->ic >= 0 ? y + (long long)S->ic * S->ng : nullptr;
   if (q && launch_quat_grad(L, p, q, nullptr, S->n2, st)) return -1;
   if (c && kks(S, y)) return -1;
   // phase block: C = 1 + gamma M w g'' (source/PhaseFACOps.cc:33-51)
   // phase mobility after the phase concentrations (QuatIntegrator::
   // setCoefficients -> computeMobilities, source/QuatIntegrator.cc:3078-3080)
   bool kim = p.phi_mobility_type == AMPE_PHI_MOB_KIM;
   if (kim && launch_phase_mobility(L, p, S->cl, S->cs, S->Mphi, st)) return -1;
   {
      // several orders: the well scale passed is 0 (source/QuatIntegrator.cc
      // :3321-3329), so C = 1
      ampe_model_params pw = p;
      if (S->nord > 1) pw.phi_well_scale = 0.0;
      if (launch_phase_C(L, pw, gamma, phi, kim ? S->Mphi : nullptr, S->Cphi, st)) return -1;
   }
   // composition block: D = -gamma D_c (source/ConcFACOps.cc:28-61); for EBS
   // D_c = D_L + D_A (+ D_B) (EBSCompositionRHSStrategy::
   // setDiffusionCoeffForPreconditionerOnPatch)
   if (c && p.conc_rhs_form == AMPE_CONC_RHS_EBS) {
      if (vec_axpby(S->nf, 1.0, S->Dc, 1.0, S->Dphi, S->Dconc, st)) return -1;
      if (S->tern && vec_axpby(S->nf, 1.0, S->Dconc, 1.0, S->Db, S->Dconc, st)) return -1;
      if (vec_axpby(S->nf, -gamma, S->Dconc, 0.0, S->Dconc, S->Dconc, st)) return -1;
   } else if (c && vec_axpby(S->nf, -gamma, S->Dc, 0.0, S->Dc, S->Dconc, st)) {
      return -1;
   }
   // orientation block (source/QuatFACOps.cc:735-829)
   if (q) {
      if (launch_quat_face_coef(L, p, phi, S->n2, S->fc, st)) return -1;
      if (launch_sqrt_quat_mobility(L, p, phi, S->sqrtm, st)) return -1;
   }
   if (S->s.precision == AMPE_FAC_FP32) {
      if (vec_to_f32(S->ng, S->Cphi, S->f_Cphi, st)) return -1;
      if (kim && vec_to_f32(S->ng, S->Mphi, S->f_Mphi, st)) return -1;
      if (c && vec_to_f32(S->nf, S->Dconc, S->f_Dconc, st)) return -1;
      if (q) {
         if (vec_to_f32(S->nf, S->fc, S->f_fc, st)) return -1;
         if (vec_to_f32(S->ng, S->sqrtm, S->f_sqrtm, st)) return -1;
      }
   }
   if (S->mgf) {
      if (mg_setup_phase<float>(*S->mgf, S->f_Cphi, kim ? S->f_Mphi : nullptr, st)) return -1;
      if (c && mg_setup_conc<float>(*S->mgf, S->f_Dconc, st)) return -1;
      if (q && mg_setup_quat<float>(*S->mgf, S->f_fc, S->f_sqrtm, st)) return -1;
   }
   if (S->mgd) {
      if (mg_setup_phase<double>(*S->mgd, S->Cphi, kim ? S->Mphi : nullptr, st)) return -1;
      if (c && mg_setup_conc<double>(*S->mgd, S->Dconc, st)) return -1;
      if (q && mg_setup_quat<double>(*S->mgd, S->fc, S->sqrtm, st)) return -1;
   }
   S->gamma_setup = gamma;
   return 0;
}

// One preconditioner block: an elliptic FAC block (phase, composition) or the
// orientation block, with its buffers.  comp = first component index.
template <typename T>
struct PBlock {
   int kind;  // 0 elliptic, 1 orientation
   int comp, ncomp;
   EllOp<T> op;
   const T* fc;
   const T* sm;
   const T* f;  // right-hand side
   T* u;        // solution
   T* e;        // cycle correction
   T* r;        // residual
   double norm;
   int cycles;
};

template <typename T>
static Mg<T>* mg_of(Sim* S);
template <>
Mg<float>* mg_of<float>(Sim* S)
{
   return S->mgf;
}
template <>
Mg<double>* mg_of<double>(Sim* S)
{
   return S->mgd;
}

template <typename T>
static int block_smooth(Sim* S, PBlock<T>& B, const T* rhs, T* x)
{
   const Level& L = *S->L;
   hipStream_t st = S->stream;
   const double* dx = S->p.dx;
   Mg<T>* mg = mg_of<T>(S);
   if (mg) {
      // one multigrid V-cycle from x = 0 (x is zero on entry)
      MgCycle cy;
      cy.nu1 = cy.nu2 = B.kind == 0 ? S->s.fac_sweeps : S->s.quat_sweeps;
      cy.ncoarse = S->s.mg_coarse_sweeps;
      if (B.kind == 0) return mg_vcycle_ell<T>(*mg, B.op, B.comp == 0 ? 0 : 1, cy, rhs, x, st);
      return mg_vcycle_quat<T>(*mg, B.fc, B.sm, S->gamma_setup, B.ncomp, cy, rhs, x, st);
   }
   if (B.kind == 0) {
      bool want = S->s.fac_tol >= 0.0;
      if (ghost_fill(L, x, 1, st)) return -1;
      for (int sw = 0; sw < S->s.fac_sweeps; sw++) {
         double mr = 0.0, mb = 0.0;
         if (ghost_fill(L, x, 1, st)) return -1;
         if (launch_rbgs<T>(L, B.op, dx, rhs, x, 0, want, st)) return -1;
         if (want && reduce_partials(L, L.ncell_tiles, 1, 1, &mr, st)) return -1;
         if (ghost_fill(L, x, 1, st)) return -1;
         if (launch_rbgs<T>(L, B.op, dx, rhs, x, 1, want, st)) return -1;
         if (want && reduce_partials(L, L.ncell_tiles, 1, 1, &mb, st)) return -1;
         if (ghost_fill(L, x, 1, st)) return -1;
         if (want && (mr > mb ? mr : mb) <= S->s.fac_tol) break;
      }
   } else {
      int ql = B.ncomp;
      if (ghost_fill(L, x, ql, st)) return -1;
      for (int sw = 0; sw < S->s.quat_sweeps; sw++) {
         if (ghost_fill(L, x, ql, st)) return -1;
         if (launch_quat_rbgs<T>(L, B.fc, B.sm, S->gamma_setup, dx, ql, rhs, x, 0, st)) return -1;
         if (ghost_fill(L, x, ql, st)) return -1;
         if (launch_quat_rbgs<T>(L, B.fc, B.sm, S->gamma_setup, dx, ql, rhs, x, 1, st)) return -1;
         if (ghost_fill(L, x, ql, st)) return -1;
      }
   }
   return 0;
}

// residual r = f - A u (ghost fill of u first), per-tile sums into region reg
template <typename T>
static int block_residual(Sim* S, PBlock<T>& B, int reg)
{
   const Level& L = *S->L;
   hipStream_t st = S->stream;
   double* part = L.partial + (size_t)reg * L.ncell_tiles;
   Mg<T>* mg = mg_of<T>(S);
   const bool split = B.kind == 1 && mg && mg->qsplit;
   if (ghost_fill(L, B.u, B.ncomp, st, split)) return -1;
   if (B.kind == 0) return launch_residual<T>(L, B.op, S->p.dx, B.f, B.u, B.r, part, st);
   return launch_quat_residual<T>(L, B.fc, B.sm, S->gamma_setup, S->p.dx, B.ncomp, B.f, B.u,
                                  B.r, part, st, split);
}

// Each block runs the single-level FAC loop of FACPreconditioner::solveSystem
// (source/samrai/FACPreconditioner.cc:267-429) with the redblack coarse solve:
// cycle 1 starts from u = 0, so its correction is smoothed directly into u
// (the residual of u = 0 is f); every cycle ends with the residual RMS.  The
// blocks' norms are reduced together, one host synchronisation per cycle.
template <typename T>
static int run_blocks(Sim* S, std::vector<PBlock<T>>& blocks, double tol)
{
   const Level& L = *S->L;
   hipStream_t st = S->stream;
   int nb = (int)blocks.size();
   bool multi = S->s.fac_max_cycles > 1 || S->s.quat_max_cycles > 1;
   // cycle 1 from u = 0 (zeroed by psolve_in)
   for (int b = 0; b < nb; b++) {
      PBlock<T>& B = blocks[b];
      if (block_smooth<T>(S, B, B.f, B.u)) return -1;
      B.cycles = 1;
      if (multi && block_residual<T>(S, B, b)) return -1;
   }
   // With one cycle per block nothing depends on the residual norms (the
   // return value is informational), so they are not computed.
   if (!multi) return 0;
   double ss[8];
   if (reduce_regions(L, L.ncell_tiles, nb, ss, st)) return -1;
   for (int b = 0; b < nb; b++)
      blocks[b].norm = std::sqrt(ss[b] / ((double)blocks[b].ncomp * L.ncell));
   for (;;) {
      std::vector<int> active;
      for (int b = 0; b < nb; b++) {
         PBlock<T>& B = blocks[b];
         int maxc = B.kind == 0 ? S->s.fac_max_cycles : S->s.quat_max_cycles;
         if (B.norm > tol && B.cycles < maxc) active.push_back(b);
      }
      if (active.empty()) break;
      for (size_t a = 0; a < active.size(); a++) {
         PBlock<T>& B = blocks[active[a]];
         if (fill_zero<T>(B.e, S->ng * B.ncomp, st)) return -1;
         if (block_smooth<T>(S, B, B.r, B.e)) return -1;
         long long n = S->ng * B.ncomp;
         if (std::is_same<T, double>::value) {
            if (vec_axpby(n, 1.0, (const double*)B.u, 1.0, (const double*)B.e, (double*)B.u, st))
               return -1;
         } else {
            if (vec_axpby_f(n, (float*)B.u, (const float*)B.e, st)) return -1;
         }
         if (block_residual<T>(S, B, (int)a)) return -1;
         B.cycles++;
      }
      if (reduce_regions(L, L.ncell_tiles, (int)active.size(), ss, st)) return -1;
      for (size_t a = 0; a < active.size(); a++) {
         PBlock<T>& B = blocks[active[a]];
         B.norm = std::sqrt(ss[a] / ((double)B.ncomp * L.ncell));
      }
   }
   int ret = 0;
   for (auto& B : blocks)
      if (!(B.norm < tol)) ret = 1;
   return ret;
}

// CVSpgmrPrecondSolve (source/QuatIntegrator.cc:3666-3771): phase block
// (PhasePrecondSolve :3391-3449), orientation block (QuatPrecondSolve
// :3594-3650, divide / multiply by sqrt(M) of QuatSysSolver::solveSystem,
// source/QuatSysSolver.cc:296-349), composition block
// (ConcentrationPrecondSolve :3547-3591).
template <typename T>
static std::vector<PBlock<T>> make_blocks(Sim* S, double dphi, const T* Cphi, const T* Mphi,
                                          const T* Dconc, const T* fc, const T* sm)
{
   const ampe_model_params& p = S->p;
   std::vector<PBlock<T>> bl;
   // one phase block per order (PhasePrecondSolve, source/QuatIntegrator.cc
   // :3427-3444)
   for (int m = 0; m < S->nord; m++) {
      PBlock<T> ph{};
      ph.kind = 0;
      ph.comp = m;
      ph.ncomp = 1;
      ph.op.Dc = (T)dphi;
      ph.op.C = Cphi;
      ph.op.Cc = T(1);
      ph.op.Mc = (T)p.phi_mobility;
      if (p.phi_mobility_type == AMPE_PHI_MOB_KIM) ph.op.M = Mphi;
      bl.push_back(ph);
   }
   if (S->iq >= 0) {
      PBlock<T> qb{};
      qb.kind = 1;
      qb.comp = S->iq;
      qb.ncomp = p.qlen;
      qb.fc = fc;
      qb.sm = sm;
      bl.push_back(qb);
   }
   // one composition block per species (ConcFACOps depth ic, source/
   // ConcFACOps.cc:38-56; the ternary diagonals are equal)
   for (int m = 0; m < S->ncs; m++) {
      PBlock<T> cb{};
      cb.kind = 0;
      cb.comp = S->ic + m;
      cb.ncomp = 1;
      cb.op.D = Dconc;
      cb.op.Cc = T(1);
      cb.op.Mc = (T)p.conc_mobility;
      bl.push_back(cb);
   }
   return bl;
}

// f_u = P^-1 f_f on the FP32 buffers (f_u zeroed first when zero_u).
static int psolve_f32(Sim* S, double delta, bool zero_u)
{
   long long ng = S->ng;
   if (zero_u && fill_zero<float>(S->f_u, S->nv, S->stream)) return -1;
   double dphi = -S->gamma_setup * S->p.epsilon_phi * S->p.epsilon_phi;
   auto bl = make_blocks<float>(S, dphi, S->f_Cphi, S->f_Mphi, S->f_Dconc, S->f_fc,
                                S->f_sqrtm);
   for (auto& B : bl) {
      long long off = (long long)B.comp * ng;
      B.f = S->f_f + off;
      B.u = S->f_u + off;
      B.e = S->f_e + off;
      B.r = S->f_r + off;
   }
   return run_blocks<float>(S, bl, delta);
}

// The solve above from a zero start has a fixed launch sequence on fixed
// buffers when every block runs one cycle (no residual norms, no host
// synchronisation: multigrid, or red-black sweeps without the early exit);
// its only launch arguments that change between solves are
// gamma's (the orientation operator and the phase block's D = -gamma eps^2).
static bool psolve_graph_ok(const Sim* S)
{
   return S->s.graph_psolve && S->s.precision == AMPE_FAC_FP32 && S->s.fac_max_cycles <= 1 &&
          S->s.quat_max_cycles <= 1 && (S->mgf || S->s.fac_tol < 0.0);
}

// psolve_f32(zero_u) captured into a HIP graph (again when gamma changed) and
// replayed: one launch instead of one per kernel of every V-cycle.  The
// coefficient buffers are read at replay, so a preconditioner setup at the
// same gamma needs no new capture.
static int psolve_f32_graph(Sim* S)
{
   hipStream_t st = S->stream;
   if (S->pgraph && S->pgraph_gamma != S->gamma_setup) {
      (void)hipGraphExecDestroy(S->pgraph);
      S->pgraph = nullptr;
   }
   if (!S->pgraph) {
      hipGraph_t g = nullptr;
      AMPE_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
      int rc = psolve_f32(S, 0.0, true);
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc < 0 || e != hipSuccess || !g) {
         if (g) (void)hipGraphDestroy(g);
         if (rc >= 0) set_error(std::string("psolve graph capture: ") + hipGetErrorString(e));
         return -1;
      }
      e = hipGraphInstantiate(&S->pgraph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
         S->pgraph = nullptr;
         set_error(std::string("psolve graph instantiate: ") + hipGetErrorString(e));
         return -1;
      }
      S->pgraph_gamma = S->gamma_setup;
   }
   AMPE_HIP_CHECK(hipGraphLaunch(S->pgraph, st));
   return 0;
}

// The level whose row parity-split layout the FP32 orientation vectors
// f_f / f_u use (null: natural layout).
static const Level* qsplit_level(const Sim* S)
{
   return S->mgf && S->mgf->qsplit ? S->L : nullptr;
}

// z = P^-1 r.  FP32 preconditioner only: f_ready = the FP32 right-hand side
// S->f_f was already formed from r (by the Krylov kernel that produced r);
// keep = the solution is left in S->f_u for the caller (z untouched).
struct PsolveOpts {
   bool f_ready = false;
   bool keep = false;
};
static int hier_psolve(Sim* S, const double* r, double* z, double delta);
static int psolve(Sim* S, const double* r, double* z, double delta, PsolveOpts o = {})
{
   if (S->hier) return hier_psolve(S, r, z, delta);
   const ampe_model_params& p = S->p;
   hipStream_t st = S->stream;
   long long ng = S->ng;
   S->st.nps++;
   double dphi = -S->gamma_setup * p.epsilon_phi * p.epsilon_phi;
   int ql = S->iq >= 0 ? p.qlen : 0;
   int iq = S->iq >= 0 ? S->iq : S->nc;
   if (S->s.precision == AMPE_FAC_FP32) {
      // f = r (orientation blocks / sqrt(M), FP64 division) in FP32, u = 0
      int ret;
      if (o.f_ready) {
         ret = psolve_graph_ok(S) ? psolve_f32_graph(S) : psolve_f32(S, delta, true);
      } else {
         if (psolve_in<float>(ng, S->nc, iq, ql, r, S->sqrtm, S->f_f, nullptr, S->f_u, st,
                              qsplit_level(S)))
            return -1;
         ret = psolve_f32(S, delta, false);
      }
      if (ret < 0) return -1;
      if (!o.keep && psolve_out<float>(ng, S->nc, iq, ql, S->f_u, S->sqrtm, nullptr, z, st,
                                       qsplit_level(S)))
         return -1;
      return ret;
   }
   if (o.f_ready || o.keep) {
      set_error("psolve: fused input/output is for the FP32 preconditioner");
      return -1;
   }
   // FP64: z = 0, fac_u = orientation blocks of r / sqrt(M)
   if (psolve_in<double>(ng, S->nc, iq, ql, r, S->sqrtm, nullptr, S->fac_u, z, st)) return -1;
   auto bl = make_blocks<double>(S, dphi, S->Cphi, S->Mphi, S->Dconc, S->fc, S->sqrtm);
   for (auto& B : bl) {
      long long off = (long long)B.comp * ng;
      B.f = B.kind == 1 ? S->fac_u : r + off;
      B.u = z + off;
      B.e = S->fac_e + off;
      B.r = S->fac_r + off;
   }
   int ret = run_blocks<double>(S, bl, delta);
   if (ret < 0) return -1;
   if (S->iq >= 0 && vec_mul_comp(S, z + (long long)S->iq * ng, S->sqrtm, ql)) return -1;
   return ret;
}

int vec_mul_comp(Sim* S, double* x, const double* sm, int ncomp)
{
   return vec_scale_comp(S->ng, ncomp, sm, x, S->stream);
}

// ------------------------------------------------------------ Krylov
// Summed control-volume weights of all components in level-0 cells: the
// cell count times the components on one level; on a hierarchy the weights
// (zero under a finer level) summed.
static double wcount(const Sim* S)
{
   return S->hier ? S->wsum : (double)S->nc * S->L->ncell;
}

static double cellvol(const Sim* S)
{
   const double* dx = S->p.dx;
   return dx[0] * dx[1] * (S->p.ndim == 3 ? dx[2] : 1.0);
}

// SPGMR, left preconditioned, MGS, x0 = 0, scaling by ewt, difference-quotient
// Jv (sigma = 1/||v||_wrms) and A v = v - gamma Jv.  V[l] is normalised lazily
// by the next iteration's first kernel.  Returns 0 converged, 1 residual
// reduced, 2 failure, -1 fatal.
static int spgmr(Sim* S, double gamma, const double* bvec, double* x, double delta)
{
   const Level& L = *S->L;
   hipStream_t st = S->stream;
   int maxl = S->s.maxl, nc = S->nc;
   double cv = cellvol(S);
   double wfac = 1.0 / wcount(S);
   std::vector<double>& Hm = S->H;
   Hm.assign((size_t)(maxl + 1) * (maxl + 1), 0.0);
   int ld = maxl + 1;
#define HH(i, j) Hm[(size_t)(i) * ld + (j)]
   double givens[2 * kMaxKrylov], g[kMaxKrylov + 1], yv[kMaxKrylov];
   // Classical Gram-Schmidt keeps the basis unnormalised: basis vector i is
   // V[i] hs[i].  The difference-quotient increment then needs no norm pass:
   // the normalised vector has sum(v^2) cv = 1, so its WRMS norm over ewt is
   // sqrt(wfac / cv) and sigma = sqrt(cv / wfac).  With the FP32
   // preconditioner, the kernel that forms A v also writes the FP32
   // preconditioner input, and the dot kernel reads the FP32 solution.
   const bool cgs = S->s.gstype == AMPE_GS_CLASSICAL;
   // (one level only: the hierarchy's FP32 block solves take the FP64 input
   // and write the FP64 output)
   const bool fuse_out = cgs && S->s.precision == AMPE_FAC_FP32 && !S->hier;
   PrecIn pin;
   PrecOut pout;
   if (fuse_out) {
      pin.f = S->f_f;
      pin.sm = S->sqrtm;
      pin.iq = S->iq >= 0 ? S->iq : nc;
      pin.ql = S->iq >= 0 ? S->p.qlen : 0;
      pout.u = S->f_u;
      pout.sm = S->sqrtm;
      pout.ewt = S->ewt;
      pout.iq = pin.iq;
      pout.ql = pin.ql;
      pin.split = pout.split = qsplit_level(S) ? 1 : 0;
   }
   double hs[kMaxKrylov + 1];
   if (psolve(S, bvec, S->work, delta) < 0) return -1;
   double rho;
   if (kry_scale_dot(L, nc, S->ewt, S->work, S->V[0], S->V[0], &rho, st)) return -1;
   rho = std::sqrt(rho * cv);
   double rho0 = rho;
   if (rho <= delta) {
      AMPE_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * S->nv, st));
      return 0;
   }
   double hinv = 1.0 / rho;
   hs[0] = hinv;
   std::memset(g, 0, sizeof(g));
   g[0] = rho;
   int lused = 0;
   bool converged = false;
   for (int l = 0; l < maxl; l++) {
      S->st.nli++;
      if (cgs) {
         double sig = std::sqrt(cv / wfac);
         if (kry_mid_scaled(L, nc, sig, S->V[l], hs[l], S->ewt, S->ycur, S->work2, st))
            return -1;
         if (rhs_eval(S, S->work2, S->jv, true)) return -1;
         double siginv = 1.0 / sig;
         // with the fused FP32 preconditioner input the FP64 (I - gamma J) v is
         // not stored: the dot kernel writes V[l+1] from the preconditioned vector
         if (kry_finish_scaled(L, nc, siginv, gamma, S->V[l], hs[l], S->ewt, S->jv, S->fcur,
                               fuse_out ? nullptr : S->V[l + 1], pin, st))
            return -1;
      } else {
         double ss;
         // V[l] *= hinv; xk = V[l] / ewt; ||xk||_wrms
         if (kry_prep(L, nc, hinv, S->V[l], S->ewt, S->xk, &ss, st)) return -1;
         double sig = 1.0 / std::sqrt(ss * wfac);
         if (kry_mid(L, nc, sig, S->xk, S->ycur, S->work2, st)) return -1;
         if (rhs_eval(S, S->work2, S->jv, true)) return -1;
         double siginv = 1.0 / sig;
         if (kry_finish(L, nc, siginv, gamma, S->xk, S->jv, S->fcur, S->V[l + 1], st))
            return -1;
      }
      if (fuse_out) {
         // input formed by the finish kernel, solution left in f_u
         PsolveOpts o;
         o.f_ready = true;
         o.keep = true;
         if (psolve(S, S->V[l + 1], nullptr, delta, o) < 0) return -1;
      } else if (psolve(S, S->V[l + 1], S->work, delta) < 0) {
         return -1;
      }
      double hn;
      if (S->s.gstype == AMPE_GS_CLASSICAL) {
         // V[l+1] = ewt work with all its dots in one pass, one linear
         // combination, re-orthogonalised when the norm dropped by > 1000
         // (SUNClassicalGS)
         VecList vl;
         vl.n = l + 1;
         for (int i = 0; i <= l; i++) vl.v[i] = S->V[i];
         double sums[kMaxCgs];
         if (fuse_out) {
            // V[l+1] = ewt P^-1 (...) from the FP32 solution, with its dots
            if (kry_cgs_dots_from(L, nc, pout, S->V[l + 1], vl, sums, st)) return -1;
         } else if (kry_cgs_dots(L, nc, S->ewt, S->work, S->V[l + 1], vl, sums, st)) {
            return -1;
         }
         double vk_norm = std::sqrt(sums[l + 1] * cv);
         for (int i = 0; i <= l; i++) {
            HH(i, l) = (sums[i] * hs[i]) * cv;
            vl.y[i] = -HH(i, l) * hs[i];
         }
         double ss;
         if (kry_cgs_update(L, nc, S->V[l + 1], vl, &ss, st)) return -1;
         hn = std::sqrt(ss * cv);
         if (1000.0 * hn < vk_norm) {
            if (kry_cgs_dots(L, nc, nullptr, nullptr, S->V[l + 1], vl, sums, st)) return -1;
            for (int i = 0; i <= l; i++) {
               double d = (sums[i] * hs[i]) * cv;
               HH(i, l) += d;
               vl.y[i] = -d * hs[i];
            }
            if (kry_cgs_update(L, nc, S->V[l + 1], vl, &ss, st)) return -1;
            hn = std::sqrt(ss * cv);
         }
      } else {
         // V[l+1] = ewt work; modified Gram-Schmidt, each step fused with the next dot
         double dot;
         if (kry_scale_dot(L, nc, S->ewt, S->work, S->V[l + 1], S->V[0], &dot, st)) return -1;
         for (int i = 0; i <= l; i++) {
            double hij = dot * cv;
            HH(i, l) = hij;
            const double* nxt = (i < l) ? S->V[i + 1] : nullptr;
            if (kry_mgs(L, nc, S->V[l + 1], S->V[i], hij, nxt, &dot, st)) return -1;
         }
         hn = std::sqrt(dot * cv);
      }
      HH(l + 1, l) = hn;
      for (int i = 0; i < l; i++) {
         double c = givens[2 * i], s = givens[2 * i + 1];
         double t1 = HH(i, l), t2 = HH(i + 1, l);
         HH(i, l) = c * t1 - s * t2;
         HH(i + 1, l) = s * t1 + c * t2;
      }
      double a = HH(l, l), bb = HH(l + 1, l), c, s;
      if (bb == 0.0) {
         c = 1.0;
         s = 0.0;
      } else if (std::fabs(bb) > std::fabs(a)) {
         double t = a / bb;
         s = -1.0 / std::sqrt(1.0 + t * t);
         c = -s * t;
      } else {
         double t = bb / a;
         c = 1.0 / std::sqrt(1.0 + t * t);
         s = -c * t;
      }
      givens[2 * l] = c;
      givens[2 * l + 1] = s;
      HH(l, l) = c * a - s * bb;
      HH(l + 1, l) = 0.0;
      g[l + 1] = s * g[l];
      g[l] = c * g[l];
      rho = std::fabs(g[l + 1]);
      lused = l + 1;
      if (rho <= delta) {
         converged = true;
         break;
      }
      if (hn == 0.0) break;
      hinv = 1.0 / hn;
      hs[l + 1] = hinv;
   }
   for (int i = lused - 1; i >= 0; i--) {
      double sacc = g[i];
      for (int j = i + 1; j < lused; j++) sacc -= HH(i, j) * yv[j];
      yv[i] = sacc / HH(i, i);
   }
   VecList vl;
   vl.n = lused;
   for (int i = 0; i < lused; i++) {
      vl.v[i] = S->V[i];
      vl.y[i] = cgs ? yv[i] * hs[i] : yv[i];
   }
   if (kry_lincomb(L, nc, vl, S->ewt, x, st)) return -1;
#undef HH
   if (converged) return 0;
   return rho < rho0 ? 1 : 2;
}

// ------------------------------------------------------------ BDF step
// tq4: the Newton convergence bound (CVODE's tq[4] = nlscoef / tq[2]; nlscoef
// in the imposed-step mode).
static int newton(Sim* S, double gamma, double delta, double tq4, bool force_setup,
                  bool* did_setup)
{
   const Level& L = *S->L;
   hipStream_t st = S->stream;
   long long nv = S->nv;
   double wfac = 1.0 / wcount(S);
   bool call_setup = force_setup || S->nst == 0 ||
                     std::fabs(gamma / S->gamma_setup - 1.0) > 0.3 ||
                     S->nst >= S->nstlp + S->s.msbp;
   *did_setup = false;
   AMPE_HIP_CHECK(hipMemcpyAsync(S->ycur, S->ypred, sizeof(double) * nv,
                                 hipMemcpyDeviceToDevice, st));
   double delp = 0.0;
   for (int m = 0; m < S->s.max_newton; m++) {
      S->st.nni++;
      if (rhs_eval(S, S->ycur, S->fcur, false)) return -1;
      if (kry_newton_rhs(L, S->nc, gamma, S->fcur, S->ycur, S->ahist, S->b, st)) return -1;
      if (m == 0 && call_setup) {
         if (psetup(S, S->ycur, gamma)) return -1;
         S->nstlp = S->nst;
         S->crate = 1.0;
         *did_setup = true;
      }
      int ls = spgmr(S, gamma, S->b, S->del, delta);
      if (ls < 0) return -1;
      if (ls == 2 || (ls == 1 && m > 0)) {
         S->st.nlcf++;
         return 1;
      }
      double ss;
      if (kry_newton_update(L, S->nc, S->del, S->ewt, S->ycur, &ss, st)) return -1;
      double del = std::sqrt(ss * wfac);
      if (m > 0) {
         double r = del / delp;
         S->crate = (0.3 * S->crate > r) ? 0.3 * S->crate : r;
      }
      double dcon = del * (S->crate < 1.0 ? S->crate : 1.0) / tq4;
      if (dcon <= 1.0) return 0;
      if (m >= 1 && del > 2.0 * delp) return 1;
      delp = del;
   }
   return 1;
}

// ScalarTemperatureStrategy::getCurrentTemperature
// (source/ScalarTemperatureStrategy.cc:56-72)
static double ramp_temperature(const Sim* S, double t)
{
   double T = S->T0 + S->dTdt * t;
   if (S->dTdt < 0.0 && T < S->Ttarget)
      T = S->Ttarget;
   else if (S->Ttarget > 0.0 && S->dTdt > 0.0 && T > S->Ttarget)
      T = S->Ttarget;
   return T;
}

int sim_set_temperature_ramp(Sim* S, double T0, double dTdt, double T_target,
                             const ampe_calphad_tdep* tab)
{
   if (!(T0 > 0.0)) {
      set_error("sim_set_temperature_ramp: T0 must be positive");
      return -1;
   }
   S->tramp = true;
   S->T0 = T0;
   S->dTdt = dTdt;
   S->Ttarget = T_target;
   S->has_tab = tab != nullptr;
   if (tab) S->tab = *tab;
   return 0;
}

static int bdf_step(Sim* S)
{
   const Level& L = *S->L;
   hipStream_t st = S->stream;
   long long nv = S->nv, ng = S->ng;
   double wfac = 1.0 / wcount(S);
   if (S->ic >= 0) {
      // resetRefPhaseConcentrations (source/QuatModel.cc:5218-5235)
      AMPE_HIP_CHECK(hipMemcpyAsync(S->clr, S->cl, sizeof(double) * ng,
                                    hipMemcpyDeviceToDevice, st));
      AMPE_HIP_CHECK(hipMemcpyAsync(S->csr, S->cs, sizeof(double) * ng,
                                    hipMemcpyDeviceToDevice, st));
   }
   double nrmfac = std::sqrt(wcount(S) * cellvol(S));
   double delta = S->s.eplifac * S->s.nlscoef * nrmfac;
   bool force_setup = false;
   for (;;) {
      double h = S->h, gamma;
      // the temperature field is set at the time of every RHS evaluation
      // (source/QuatIntegrator.cc:3191); a step attempt evaluates at t + h
      if (S->tramp && S->dTdt != 0.0) {
         S->p.temperature = ramp_temperature(S, S->t + h);
         if (S->has_tab) calphad_at(S->tab, S->p.temperature, S->p.calphad);
      }
      if (!S->have_prev) {
         AMPE_HIP_CHECK(hipMemcpyAsync(S->ypred, S->y, sizeof(double) * nv,
                                       hipMemcpyDeviceToDevice, st));
      } else {
         double w = h / S->hprev;
         // ypred = y + w (y - yprev)
         if (vec_axpby(nv, 1.0, S->y, -1.0, S->yprev, S->work, st)) return -1;
         if (vec_axpby(nv, 1.0, S->y, w, S->work, S->ypred, st)) return -1;
      }
      if (!S->have_prev || S->s.max_order < 2) {
         gamma = h;
         AMPE_HIP_CHECK(hipMemcpyAsync(S->ahist, S->y, sizeof(double) * nv,
                                       hipMemcpyDeviceToDevice, st));
      } else {
         double w = h / S->hprev;
         double den = 1.0 + 2.0 * w;
         gamma = h * (1.0 + w) / den;
         double c1 = (1.0 + w) * (1.0 + w) / den, c2 = w * w / den;
         if (vec_bdf2_hist(nv, c1, S->y, c2, S->yprev, S->ahist, st)) return -1;
      }
      bool did_setup = false;
      int r = newton(S, gamma, delta, S->s.nlscoef, force_setup, &did_setup);
      if (r < 0) return -1;
      if (kks_collect(S)) return -1;
      if (r == 0) {
         if (vec_axpby(nv, 1.0, S->ycur, -1.0, S->ypred, S->acor, st)) return -1;
         if (S->iq >= 0 && S->p.qlen > 1) {
            AMPE_HIP_CHECK(hipMemsetAsync(S->corr, 0, sizeof(double) * nv, st));
            if (launch_quat_project(L, S->p, S->ycur + (long long)S->iq * ng,
                                    S->corr + (long long)S->iq * ng,
                                    S->acor + (long long)S->iq * ng, st))
               return -1;
            if (vec_axpby(nv, 1.0, S->ycur, 1.0, S->corr, S->ycur, st)) return -1;
         }
         double ss;
         if (vec_sumsq_w(L, S->nc, S->acor, S->ewt, &ss, st)) return -1;
         S->st.last_acnrm = std::sqrt(ss * wfac);
         std::swap(S->yprev, S->y);
         std::swap(S->y, S->ycur);
         S->have_prev = true;
         S->hprev = h;
         S->t += h;
         S->nst++;
         S->st.nst++;
         S->h = (2.0 * h < S->s.dt) ? 2.0 * h : S->s.dt;
         if (vec_ewt(nv, S->y, S->s.rtol, S->s.atol, S->ewt, st)) return -1;
         return 0;
      }
      S->st.ncfn++;
      if (!did_setup) {
         force_setup = true;
         continue;
      }
      S->h = 0.25 * h;
      force_setup = true;
      if (S->h < S->s.dt * 1.0e-8) return 1;
   }
}

// ------------------------------------------------------------ adaptive BDF
// CVODE's BDF as QuatIntegrator configures it (source/QuatIntegrator.cc
// :1571-1596; SUNDIALS CVODE 5.x, orders 1..2): the Nordsieck history
// z = (y, h y', h^2 y''/2) is predicted by the Pascal matrix, the corrector
// solves y_n = y_pred + acor with acor = gamma f(y_n) - z1 / l1, gamma = h / l1,
// the local error test is ||acor||_wrms tq2 <= 1, and after each step the next
// step and order follow from eta_q = 1 / ((6 dsm)^(1/(q+1)) + 1e-6) and its
// neighbours of order q - 1 (6 ||z_q|| tq1) and q + 1 (10 ||acor - c z_saved||
// tq3), changed only when eta >= 1.5.  Error-test failures cut h by
// max(0.1, eta_q) (at most 0.2 from the second failure; after three, order
// one and h / 10 with a fresh history); convergence failures by 4.

// Method and error-test coefficients of order q for the step h (CVODE's
// fixed-leading-coefficient BDF with tau[1] the previous step).
static void bdf_coeffs(BdfCtl& B, double h, double nlscoef)
{
   const int q = B.q;
   double l0 = 1.0, l1 = 1.0, l2 = 0.0;
   double alpha0 = -1.0, alpha0_hat = -1.0, xi_inv = 1.0, xistar_inv = 1.0;
   double hsum = h;
   if (q == 2) {
      alpha0 -= 1.0 / 2.0;
      xistar_inv = -l1 - alpha0;
      hsum += B.tau[1];
      xi_inv = h / hsum;
      alpha0_hat = -l1 - xi_inv;
      l2 += l1 * xistar_inv;
      l1 += l0 * xistar_inv;
   }
   B.l[0] = l0;
   B.l[1] = l1;
   B.l[2] = l2;
   const double lq = (q == 2) ? l2 : l1;
   double A1 = 1.0 - alpha0_hat + alpha0;
   double A2 = 1.0 + q * A1;
   B.tq[2] = std::fabs(A1 / (alpha0 * A2));
   B.tq[5] = std::fabs(A2 * xistar_inv / (lq * xi_inv));
   if (B.qwait == 1) {
      if (q > 1) {
         double Cc = xistar_inv / lq;
         double A3 = alpha0 + 1.0 / q;
         double A4 = alpha0_hat + xi_inv;
         double Cpinv = (1.0 - A4 + A3) / A3;
         B.tq[1] = std::fabs(Cc * Cpinv);
      } else {
         B.tq[1] = 1.0;
      }
      hsum += B.tau[q];
      xi_inv = h / hsum;
      double A5 = alpha0 - 1.0 / (q + 1);
      double A6 = alpha0_hat - xi_inv;
      double Cppinv = (1.0 - A6 + A5) / A2;
      B.tq[3] = std::fabs(Cppinv / (xi_inv * (q + 2) * A5));
   }
   B.tq[4] = nlscoef / B.tq[2];
}

static double bdf_eta_of(double bias, double err, int p)
{
   return 1.0 / (std::pow(bias * err, 1.0 / p) + 1.0e-6);
}

// z1, z2 of the history (the order-q columns; z0 is S->y)
static double* zcol(Sim* S, int j) { return j == 0 ? S->y : (j == 1 ? S->z1 : S->z2); }

// Pascal-matrix prediction of the history in place, and its inverse.
static int bdf_predict(Sim* S, double sign)
{
   const int q = S->bdf.q;
   for (int k = 1; k <= q; k++)
      for (int j = q; j >= k; j--)
         if (vec_axpby(S->nv, 1.0, zcol(S, j - 1), sign, zcol(S, j), zcol(S, j - 1), S->stream))
            return -1;
   return 0;
}

// Column j of the history times eta^j, h = hscale eta.
static int bdf_rescale(Sim* S, double eta)
{
   BdfCtl& B = S->bdf;
   double factor = eta;
   for (int j = 1; j <= B.q; j++) {
      if (vec_axpby(S->nv, factor, zcol(S, j), 0.0, zcol(S, j), zcol(S, j), S->stream))
         return -1;
      factor *= eta;
   }
   S->h = B.hscale * eta;
   B.hscale = S->h;
   return 0;
}

// WRMS norm of a state vector against the current error weights.
static int bdf_wrms(Sim* S, const double* x, double* out)
{
   double ss;
   if (vec_sumsq_w(*S->L, S->nc, x, S->ewt, &ss, S->stream)) return -1;
   *out = std::sqrt(ss / wcount(S));
   return 0;
}

static void bdf_set_eta(BdfCtl& B, double h, double hmax)
{
   if (B.eta < 1.5) {
      B.eta = 1.0;
      B.hprime = h;
   } else {
      B.eta = std::min(B.eta, B.etamax);
      B.eta /= std::max(1.0, std::fabs(h) * (1.0 / hmax) * B.eta);
      B.hprime = h * B.eta;
   }
}

// Step and order for the next step after an accepted one (acor: the step's
// correction).
static int bdf_prepare_next(Sim* S, double dsm, const double* acor, int qmax)
{
   BdfCtl& B = S->bdf;
   const double h = S->h;
   if (B.etamax == 1.0) {
      B.qwait = std::max(B.qwait, 2);
      B.qprime = B.q;
      B.hprime = h;
      B.eta = 1.0;
      return 0;
   }
   double etaq = bdf_eta_of(6.0, dsm, B.L);
   if (B.qwait != 0) {
      B.eta = etaq;
      B.qprime = B.q;
      bdf_set_eta(B, h, S->s.dt);
      return 0;
   }
   B.qwait = 2;
   double etaqm1 = 0.0, etaqp1 = 0.0;
   if (B.q > 1) {
      double n;
      if (bdf_wrms(S, zcol(S, B.q), &n)) return -1;
      etaqm1 = bdf_eta_of(6.0, n * B.tq[1], B.q);
   }
   if (B.q != qmax && B.saved_tq5 != 0.0) {
      double ratio = h / B.tau[2], pw = 1.0;
      for (int i = 0; i < B.L; i++) pw *= ratio;
      double cquot = (B.tq[5] / B.saved_tq5) * pw;
      if (vec_axpby(S->nv, -cquot, S->z2, 1.0, acor, S->work2, S->stream)) return -1;
      double n;
      if (bdf_wrms(S, S->work2, &n)) return -1;
      etaqp1 = bdf_eta_of(10.0, n * B.tq[3], B.L + 1);
   }
   double etam = std::max(etaqm1, std::max(etaq, etaqp1));
   if (etam < 1.5) {
      B.eta = 1.0;
      B.qprime = B.q;
   } else if (etam == etaq) {
      B.eta = etaq;
      B.qprime = B.q;
   } else if (etam == etaqm1) {
      B.eta = etaqm1;
      B.qprime = B.q - 1;
   } else {
      B.eta = etaqp1;
      B.qprime = B.q + 1;
      AMPE_HIP_CHECK(hipMemcpyAsync(S->z2, acor, sizeof(double) * S->nv,
                                    hipMemcpyDeviceToDevice, S->stream));
   }
   bdf_set_eta(B, h, S->s.dt);
   return 0;
}

// Order change of the history (CVODE's cvAdjustOrder for BDF with q <= 2:
// raising 1 -> 2 starts the new column at zero, lowering 2 -> 1 drops it).
static int bdf_adjust_order(Sim* S, int qnew)
{
   BdfCtl& B = S->bdf;
   if (qnew > B.q && vec_fill(S->nv, 0.0, S->z2, S->stream)) return -1;
   B.q = qnew;
   B.L = qnew + 1;
   B.qwait = B.L;
   return 0;
}

// z1 = h f(t, y) (the initial history, and its reload after repeated error
// test failures at order one)
static int bdf_load_z1(Sim* S, double t)
{
   if (S->tramp && S->dTdt != 0.0) {
      S->p.temperature = ramp_temperature(S, t);
      if (S->has_tab) calphad_at(S->tab, S->p.temperature, S->p.calphad);
   }
   if (rhs_eval(S, S->y, S->fcur, false)) return -1;
   return vec_axpby(S->nv, S->h, S->fcur, 0.0, S->fcur, S->z1, S->stream);
}

static int bdf_adaptive_step(Sim* S)
{
   const Level& L = *S->L;
   hipStream_t st = S->stream;
   long long nv = S->nv, ng = S->ng;
   BdfCtl& B = S->bdf;
   const int qmax = S->s.max_order;
   const double hmin = S->s.dt * 1.0e-8;
   if (S->ic >= 0) {
      // resetRefPhaseConcentrations (source/QuatModel.cc:5218-5235)
      AMPE_HIP_CHECK(hipMemcpyAsync(S->clr, S->cl, sizeof(double) * ng,
                                    hipMemcpyDeviceToDevice, st));
      AMPE_HIP_CHECK(hipMemcpyAsync(S->csr, S->cs, sizeof(double) * ng,
                                    hipMemcpyDeviceToDevice, st));
   }
   if (!B.started) {
      B = BdfCtl{};
      B.hscale = S->h;
      B.hprime = S->h;
      if (bdf_load_z1(S, S->t)) return -1;
      B.started = true;
   } else if (B.hprime != S->h) {
      if (B.qprime != B.q && bdf_adjust_order(S, B.qprime)) return -1;
      if (bdf_rescale(S, B.eta)) return -1;
   }
   double nrmfac = std::sqrt(wcount(S) * cellvol(S));
   int nef = 0, ncf = 0;
   bool force_setup = false, predicted = false;
   double gamma = 0.0, delta = 0.0;
   for (;;) {
      const double h = S->h;
      if (!predicted) {
         if (S->tramp && S->dTdt != 0.0) {
            S->p.temperature = ramp_temperature(S, S->t + h);
            if (S->has_tab) calphad_at(S->tab, S->p.temperature, S->p.calphad);
         }
         if (bdf_predict(S, 1.0)) return -1;
         AMPE_HIP_CHECK(hipMemcpyAsync(S->ypred, S->y, sizeof(double) * nv,
                                       hipMemcpyDeviceToDevice, st));
         bdf_coeffs(B, h, S->s.nlscoef);
         gamma = h / B.l[1];
         // the corrector y - y_pred + z1 / l1 - gamma f(y) = 0 in the form
         // y - a - gamma f(y) = 0
         if (vec_axpby(nv, 1.0, S->ypred, -1.0 / B.l[1], S->z1, S->ahist, st)) return -1;
         delta = S->s.eplifac * B.tq[4] * nrmfac;
         predicted = true;
      }
      bool did_setup = false;
      int r = newton(S, gamma, delta, B.tq[4], force_setup, &did_setup);
      if (r < 0) return -1;
      if (kks_collect(S)) return -1;
      if (r != 0) {
         S->st.ncfn++;
         if (!did_setup) {
            // retry this prediction with a fresh preconditioner
            force_setup = true;
            continue;
         }
         if (bdf_predict(S, -1.0)) return -1;
         predicted = false;
         ncf++;
         B.etamax = 1.0;
         if (ncf >= 10 || h <= hmin) return 1;
         if (bdf_rescale(S, 0.25)) return -1;
         force_setup = true;
         continue;
      }
      // converged: acor = y - y_pred, projected onto |q| = 1 with the error
      if (vec_axpby(nv, 1.0, S->ycur, -1.0, S->ypred, S->acor, st)) return -1;
      if (S->iq >= 0 && S->p.qlen > 1) {
         AMPE_HIP_CHECK(hipMemsetAsync(S->corr, 0, sizeof(double) * nv, st));
         if (launch_quat_project(L, S->p, S->ycur + (long long)S->iq * ng,
                                 S->corr + (long long)S->iq * ng,
                                 S->acor + (long long)S->iq * ng, st))
            return -1;
         if (vec_axpby(nv, 1.0, S->ycur, 1.0, S->corr, S->ycur, st)) return -1;
      }
      double acnrm;
      if (bdf_wrms(S, S->acor, &acnrm)) return -1;
      S->st.last_acnrm = acnrm;
      const double dsm = acnrm * B.tq[2];
      if (dsm > 1.0) {
         if (bdf_predict(S, -1.0)) return -1;
         predicted = false;
         nef++;
         S->st.netf++;
         B.etamax = 1.0;
         if (nef >= 7 || h <= hmin) return 1;
         if (nef <= 3) {
            double eta = std::max(0.1, bdf_eta_of(6.0, dsm, B.L));
            if (nef >= 2) eta = std::min(eta, 0.2);
            if (bdf_rescale(S, eta)) return -1;
         } else if (B.q > 1) {
            if (bdf_adjust_order(S, B.q - 1)) return -1;
            if (bdf_rescale(S, 0.1)) return -1;
         } else {
            S->h = 0.1 * h;
            B.hscale = S->h;
            B.qwait = 10;
            if (bdf_load_z1(S, S->t)) return -1;
         }
         continue;
      }
      // accepted: the history takes the projected correction y_n - y_pred
      if (vec_axpby(nv, 1.0, S->ycur, -1.0, S->ypred, S->work, st)) return -1;
      S->nst++;
      S->st.nst++;
      for (int i = B.q; i >= 2; i--) B.tau[i] = B.tau[i - 1];
      if (B.q == 1 && S->nst > 1) B.tau[2] = B.tau[1];
      B.tau[1] = h;
      for (int j = 0; j <= B.q; j++)
         if (vec_axpby(nv, B.l[j], S->work, 1.0, zcol(S, j), zcol(S, j), st)) return -1;
      B.qwait--;
      if (B.qwait == 1 && B.q != qmax) {
         AMPE_HIP_CHECK(hipMemcpyAsync(S->z2, S->work, sizeof(double) * nv,
                                       hipMemcpyDeviceToDevice, st));
         B.saved_tq5 = B.tq[5];
      }
      S->t += h;
      S->st.q = B.q;
      if (bdf_prepare_next(S, dsm, S->work, qmax)) return -1;
      B.etamax = 10.0;
      if (vec_ewt(nv, S->y, S->s.rtol, S->s.atol, S->ewt, st)) return -1;
      return 0;
   }
}

// ------------------------------------------------------------ callbacks
// The integrator callbacks on caller-owned state vectors (nc ghosted
// component blocks), for a host that keeps its own time integrator.

static void set_time(Sim* S, double t)
{
   if (S->tramp && S->dTdt != 0.0) {
      S->p.temperature = ramp_temperature(S, t);
      if (S->has_tab) calphad_at(S->tab, S->p.temperature, S->p.calphad);
   }
}

int sim_cb_rhs(Sim* S, double t, double* y, double* ydot, int fd)
{
   set_time(S, t);
   return rhs_eval(S, y, ydot, fd != 0);
}

int sim_cb_psetup(Sim* S, double t, double* y, double gamma)
{
   set_time(S, t);
   if (psetup(S, y, gamma)) return -1;
   if (S->pgraph && S->pgraph_gamma != gamma) {
      (void)hipGraphExecDestroy(S->pgraph);
      S->pgraph = nullptr;
   }
   return 0;
}

int sim_cb_psolve(Sim* S, const double* r, double* z, double delta)
{
   return psolve(S, r, z, delta);
}

int sim_cb_project(Sim* S, const double* y, double* corr, double* err)
{
   hipStream_t st = S->stream;
   AMPE_HIP_CHECK(hipMemsetAsync(corr, 0, sizeof(double) * S->nv, st));
   if (S->iq < 0 || S->p.qlen <= 1) return 0;
   long long off = (long long)S->iq * S->ng;
   return launch_quat_project(*S->L, S->p, y + off, corr + off, err + off, st);
}

int sim_cb_reset_ref(Sim* S)
{
   if (S->ic < 0) return 0;
   AMPE_HIP_CHECK(hipMemcpyAsync(S->clr, S->cl, sizeof(double) * S->ng,
                                 hipMemcpyDeviceToDevice, S->stream));
   AMPE_HIP_CHECK(hipMemcpyAsync(S->csr, S->cs, sizeof(double) * S->ng,
                                 hipMemcpyDeviceToDevice, S->stream));
   return 0;
}

int sim_cb_kks_check(Sim* S) { return kks_collect(S); }

int sim_advance(Sim* S, int nsteps, ampe_stats* stats)
{
   int ret = 0;
   AMPE_HIP_CHECK(hipEventRecord(S->ev0, S->stream));
   for (int n = 0; n < nsteps; n++) {
      ret = S->s.adaptive ? bdf_adaptive_step(S) : bdf_step(S);
      if (ret) break;
   }
   AMPE_HIP_CHECK(hipEventRecord(S->ev1, S->stream));
   AMPE_HIP_CHECK(hipEventSynchronize(S->ev1));
   float ms = 0.f;
   AMPE_HIP_CHECK(hipEventElapsedTime(&ms, S->ev0, S->ev1));
   S->last_ms = ms;
   S->st.t = S->t;
   S->st.h = S->s.adaptive ? S->bdf.hprime : S->h;
   if (!S->s.adaptive) S->st.q = S->have_prev && S->s.max_order >= 2 ? 2 : 1;
   if (stats) *stats = S->st;
   return ret;
}

}  // namespace ampe
