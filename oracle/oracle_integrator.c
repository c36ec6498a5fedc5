/*
 * Oracle integrator: FP64 CPU restatement of QuatIntegrator's CVODE loop on
 * one patch, periodic or with zero-slope boundaries.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 * Follows:
 *   - RHS callback QuatIntegrator::evaluateRHSFunction
 *     (source/QuatIntegrator.cc:3134-3298) with setCoefficients (:2994-3083),
 *     lagged quaternion side gradients in Jv evaluations (:3189, :2805-2811)
 *     and composition coefficients refreshed only when not lagged (:3245).
 *   - Preconditioner setup CVSpgmrPrecondSet (:3300-3376) and solve
 *     CVSpgmrPrecondSolve (:3666-3771): phase, orientation and composition
 *     blocks; single-level FAC = FACPreconditioner::solveSystem
 *     (source/samrai/FACPreconditioner.cc:267-429) with the "redblack"
 *     coarse solve (source/EllipticFACOps.cc:1356-1385).  The orientation
 *     block's hypre PCG/PFMG level solve (source/QuatLevelSolver.cc:1210-1339)
 *     is replaced by red-black Gauss-Seidel on the same symmetrised matrix
 *     (set_j_ij / set_stencil, source/fortran/2d/quatlevelsolver.m4:9-120):
 *     parity unpinned against hypre.
 *   - CVODE semantics as configured in source/samrai/CVODESolver.cc:150-247
 *     and source/QuatIntegrator.cc:1571-1596: Newton-Krylov with SPGMR (MGS,
 *     left preconditioning, ewt scaling, maxl, no restarts), difference
 *     quotient Jv with sigma = 1/||v||_wrms, lsetup frequency msbp,
 *     gamma-ratio test 0.3, Newton test del*min(1,crate) <= nlscoef,
 *     linear tolerance eplifac*nlscoef*sqrt(sum cv), projection of q after
 *     each step (applyProjection :3911-3965).  Step size and order are fixed
 *     (BDF1 for the first step, BDF2 after): CVODE's adaptive heuristics are
 *     parity unpinned.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

static int dbg(void)
{
   static int v = -1;
   if (v < 0) {
      const char* e = getenv("ORACLE_DEBUG");
      v = (e && e[0] && e[0] != '0') ? 1 : 0;
   }
   return v;
}

struct OSim {
   ampe_model_params p;
   ampe_solver_params s;
   OGeom G;
   int nc, iq, ic;
   int nord, ncs, tern; /* phase orders, species, three-phase ternary model */
   long nv, nf;
   double t, h, hprev, gamma_setup, crate;
   long nst, nstlp;
   int have_prev;
   double *y, *yprev, *ypred, *ycur, *ahist, *fcur, *b, *del, *ewt, *work,
       *work2, *jv, *xk, *corr, *acor;
   double** V;
   double *cl, *cs, *clr, *csr;
   double* ph6; /* ternary three-phase concentrations cL0, cL1, cA0, cA1, cB0, cB1 */
   double *n2, *Dc, *Dphi, *Db;
   double *Cphi, *Dconc, *fc, *sqrtm, *mq, *Mphi, *phid;
   double *fac_u, *fac_e, *fac_r, *fac_flux;
   double* H;
   ampe_stats st;
   int fail;
   /* adaptive BDF (s.adaptive): Nordsieck columns h y', h^2 y''/2 and the
    * step / order controller (see bdf_adaptive_step) */
   double *z1, *z2;
   int q, qprime, L, qwait, started;
   double tau[4], l[3], tq[6], hscale, hprime, eta, etamax, saved_tq5;
   /* patch hierarchy (o_sim_create_hier): composite vectors, weights */
   OHier* Hh;
   double* wt;
   double wsum;
   unsigned char* interior;
   /* scalar temperature ramp (o_sim_set_temperature_ramp) */
   int tramp, has_tab;
   double T0, dTdt, Ttarget;
   ampe_calphad_tdep tab;
};

void o_sim_set_periodic(OSim* S, int periodic) { S->G.periodic = periodic ? 1 : 0; }

void o_sim_set_temperature_ramp(OSim* S, double T0, double dTdt, double T_target,
                                const ampe_calphad_tdep* tab)
{
   S->tramp = 1;
   S->T0 = T0;
   S->dTdt = dTdt;
   S->Ttarget = T_target;
   S->has_tab = tab != NULL;
   if (tab) S->tab = *tab;
}

/* ScalarTemperatureStrategy::getCurrentTemperature
 * (source/ScalarTemperatureStrategy.cc:56-72) */
static double ramp_temperature(const OSim* S, double t)
{
   double T = S->T0 + S->dTdt * t;
   if (S->dTdt < 0.0 && T < S->Ttarget)
      T = S->Ttarget;
   else if (S->Ttarget > 0.0 && S->dTdt > 0.0 && T > S->Ttarget)
      T = S->Ttarget;
   return T;
}

static double* vec(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

/* first step: dt_init when 0 < dt_init < dt, else dt */
static double first_step(const ampe_solver_params* s)
{
   return (s->dt_init > 0.0 && s->dt_init < s->dt) ? s->dt_init : s->dt;
}

OSim* o_sim_create(const ampe_model_params* p, const ampe_solver_params* s,
                   int nx, int ny, int nz)
{
   OSim* S = (OSim*)calloc(1, sizeof(OSim));
   S->p = *p;
   S->s = *s;
   /* two ghost cells for the fourth-order composition stencil
    * (QuatModelParameters::nghosts_required, source/QuatModelParameters.h:305-311) */
   int ghost = (p->with_conc && p->conc_rhs_form == AMPE_CONC_RHS_EBS &&
                p->ebs_stencil == AMPE_EBS_ORDER4)
                   ? 2
                   : 1;
   o_geom_init(&S->G, p->ndim, nx, ny, nz, ghost, p->dx);
   /* norderp > 1: the phase field has norderp order parameters
    * (MultiOrderRHSStrategy); the three-phase ternary model carries two
    * species */
   S->nord = p->norderp > 1 ? p->norderp : 1;
   S->tern = p->fe_model == AMPE_FE_QUADRATIC_TERNARY_3PH;
   S->ncs = p->with_conc ? (S->tern ? 2 : 1) : 0;
   if ((S->tern && (S->nord < 3 || !p->with_conc || p->evolve_quat ||
                    p->conc_rhs_form != AMPE_CONC_RHS_EBS)) ||
       (S->nord > 1 && !S->tern && (p->with_conc || p->evolve_quat))) {
      free(S);
      return NULL;
   }
   S->nc = S->nord + S->ncs + (p->evolve_quat ? p->qlen : 0);
   S->ic = p->with_conc ? S->nord : -1;
   S->iq = p->evolve_quat ? S->nord + S->ncs : -1;
   S->nv = (long)S->nc * S->G.ncg;
   S->nf = S->G.nfx + S->G.nfy + S->G.nfz;
   double** vs[] = {&S->y, &S->yprev, &S->ypred, &S->ycur, &S->ahist, &S->fcur,
                    &S->b, &S->del, &S->ewt, &S->work, &S->work2, &S->jv,
                    &S->xk, &S->corr, &S->acor};
   for (size_t i = 0; i < sizeof(vs) / sizeof(vs[0]); i++) *vs[i] = vec(S->nv);
   S->V = (double**)calloc((size_t)s->maxl + 1, sizeof(double*));
   for (int i = 0; i <= s->maxl; i++) S->V[i] = vec(S->nv);
   long ng = S->G.ncg;
   S->cl = vec(ng);
   S->cs = vec(ng);
   S->clr = vec(ng);
   S->csr = vec(ng);
   S->Cphi = vec(ng);
   S->Mphi = vec(ng);
   S->phid = vec(ng);
   S->sqrtm = vec(ng);
   S->mq = vec(ng);
   S->fac_u = vec(ng * (p->evolve_quat ? p->qlen : 1));
   S->fac_e = vec(ng * (p->evolve_quat ? p->qlen : 1));
   S->fac_r = vec(ng * (p->evolve_quat ? p->qlen : 1));
   S->n2 = vec(S->nf);
   S->Dc = vec(S->nf);
   S->Dphi = vec(S->nf);
   S->Db = vec(S->nf);
   S->ph6 = vec(6 * ng);
   S->Dconc = vec(S->nf);
   S->fc = vec(S->nf);
   S->fac_flux = vec(S->nf);
   S->H = vec((long)(s->maxl + 1) * (s->maxl + 1));
   S->z1 = vec(S->nv);
   S->z2 = vec(S->nv);
   S->h = first_step(s);
   return S;
}

void o_sim_destroy(OSim* S)
{
   if (!S) return;
   double* fs[] = {S->y, S->yprev, S->ypred, S->ycur, S->ahist, S->fcur, S->b,
                   S->del, S->ewt, S->work, S->work2, S->jv, S->xk, S->corr,
                   S->acor, S->cl, S->cs, S->clr, S->csr, S->Cphi, S->Mphi, S->sqrtm,
                   S->mq, S->fac_u, S->fac_e, S->fac_r, S->n2, S->Dc, S->Dphi,
                   S->Dconc, S->fc, S->fac_flux, S->H, S->phid, S->z1, S->z2, S->Db, S->ph6};
   for (size_t i = 0; i < sizeof(fs) / sizeof(fs[0]); i++) free(fs[i]);
   for (int i = 0; i <= S->s.maxl; i++) free(S->V[i]);
   free(S->V);
   free(S->wt);
   free(S->interior);
   free(S);
}

#define CIG(G, i, j, k) \
   ((long)((i) + (G)->g) + (long)((j) + (G)->g) * (G)->sy + (long)((k) + (G)->gz) * (G)->sz)

static void load_field(const OGeom* G, double* dst, const double* src)
{
   long n = 0;
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) dst[CIG(G, i, j, k)] = src[n++];
}

static void store_field(const OGeom* G, double* dst, const double* src)
{
   long n = 0;
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) dst[n++] = src[CIG(G, i, j, k)];
}

static void update_ewt(OSim* S, const double* y)
{
   for (long n = 0; n < S->nv; n++)
      S->ewt[n] = 1.0 / (S->s.rtol * fabs(y[n]) + S->s.atol);
}

void o_sim_set_state(OSim* S, const double* phi, const double* c,
                     const double* q)
{
   const OGeom* G = &S->G;
   long N = (long)G->nx * G->ny * G->nz;
   memset(S->y, 0, sizeof(double) * S->nv);
   for (int m = 0; m < S->nord; m++) load_field(G, S->y + (long)m * G->ncg, phi + (long)m * N);
   for (int m = 0; m < S->ncs; m++)
      load_field(G, S->y + (long)(S->ic + m) * G->ncg, c + (long)m * N);
   if (S->iq >= 0)
      for (int m = 0; m < S->p.qlen; m++)
         load_field(G, S->y + (long)(S->iq + m) * G->ncg,
                    q + (long)m * G->nx * G->ny * G->nz);
   o_ghost_fill(G, S->y, S->nc);
   if (S->tern) {
      o_kks_tern3(G, &S->p, S->nord, S->y, S->y + (long)S->ic * G->ncg, S->ph6);
   } else if (S->ic >= 0) {
      /* initial phase concentrations: KKS from the guess cL = cS = c */
      const double* cc = S->y + (long)S->ic * G->ncg;
      memcpy(S->clr, cc, sizeof(double) * G->ncg);
      memcpy(S->csr, cc, sizeof(double) * G->ncg);
      o_kks_solve(G, &S->p, S->y, cc, S->clr, S->csr, S->cl, S->cs);
   }
   S->t = 0.0;
   S->nst = 0;
   S->have_prev = 0;
   S->h = first_step(&S->s);
   S->started = 0;
   memset(&S->st, 0, sizeof(S->st));
   update_ewt(S, S->y);
}

void o_sim_get_state(OSim* S, double* phi, double* c, double* q, double* cl,
                     double* cs)
{
   const OGeom* G = &S->G;
   long N = (long)G->nx * G->ny * G->nz;
   if (phi)
      for (int m = 0; m < S->nord; m++) store_field(G, phi + (long)m * N, S->y + (long)m * G->ncg);
   if (c)
      for (int m = 0; m < S->ncs; m++)
         store_field(G, c + (long)m * N, S->y + (long)(S->ic + m) * G->ncg);
   if (q && S->iq >= 0)
      for (int m = 0; m < S->p.qlen; m++)
         store_field(G, q + (long)m * N, S->y + (long)(S->iq + m) * G->ncg);
   /* the three-phase model: cl = liquid, cs = solid A, two species each */
   for (int m = 0; m < (S->tern ? 2 : 1) && S->ic >= 0; m++) {
      if (cl) store_field(G, cl + (long)m * N, S->tern ? S->ph6 + (long)m * G->ncg : S->cl);
      if (cs) store_field(G, cs + (long)m * N, S->tern ? S->ph6 + (long)(2 + m) * G->ncg : S->cs);
   }
}

/* evaluateRHSFunction (source/QuatIntegrator.cc:3134-3298).  y is ghost
 * filled in place (the scratch copy of fillScratch is the vector itself). */
static int rhs_eval_hier(OSim* S, double* y, double* ydot, int fd);
static int psetup_hier(OSim* S, double* y, double gamma);
static int psolve_hier(OSim* S, const double* r, double* z, double delta);
static void project_q(OSim* S, const double* q, double* corr, double* err);

static int rhs_eval(OSim* S, double* y, double* ydot, int fd)
{
   if (S->Hh) return rhs_eval_hier(S, y, ydot, fd);
   const OGeom* G = &S->G;
   const ampe_model_params* p = &S->p;
   long ng = G->ncg;
   int recompute = (!fd) || (!S->s.lag_quat_sidegrad);
   S->st.nfe++;
   o_ghost_fill(G, y, S->nc);
   const double* phi = y;
   const double* q = (S->iq >= 0) ? y + (long)S->iq * ng : NULL;
   const double* c = (S->ic >= 0) ? y + (long)S->ic * ng : NULL;
   if (S->tern) {
      /* three phases, two species: closed-form KKS, multi-order phase RHS with
       * the driving force, EBS composition RHS of the three phases */
      o_kks_tern3(G, p, S->nord, phi, c, S->ph6);
      S->st.kks_calls++;
      S->st.kks_cells += ng;
      o_phase_rhs_tern3(G, p, S->nord, phi, S->ph6, NULL, ydot);
      if (recompute) o_ebs3_coeffs(G, p, S->nord, phi, S->Dc, S->Dphi, S->Db);
      o_conc_rhs_ebs3(G, p, S->ph6, S->Dc, S->Dphi, S->Db, ydot + (long)S->ic * ng);
      return 0;
   }
   if (S->nord > 1) {
      o_phase_rhs_multiorder(G, p, S->nord, phi, NULL, ydot);
      return 0;
   }
   if (q && recompute) o_quat_face_norm2(G, p, q, S->n2);
   if (c) {
      int it = o_kks_solve(G, p, phi, c, S->clr, S->csr, S->cl, S->cs);
      S->st.kks_calls++;
      S->st.kks_cells += ng;
      if (it < 0) return -1;
      S->st.kks_iters += it;
   }
   o_phase_rhs(G, p, phi, q, S->cl, S->cs, ydot);
   if (q) o_quat_rhs(G, p, phi, q, S->n2, ydot + (long)S->iq * ng);
   if (c) {
      if (p->conc_rhs_form == AMPE_CONC_RHS_EBS) {
         /* Dc, Dphi hold the liquid and solid face diffusivities */
         if (recompute) {
            if (p->conc_diffusion_type == AMPE_CONC_DIFF_COMPOSITION)
               o_ebs_coeffs_mob(G, p, phi, S->cl, S->cs, S->Dc, S->Dphi);
            else
               o_ebs_coeffs(G, p, phi, S->Dc, S->Dphi);
         }
         if (p->ebs_stencil == AMPE_EBS_ISOTROPIC) {
            /* diffusivities from the phase field of the last full evaluation */
            if (recompute) memcpy(S->phid, phi, sizeof(double) * ng);
            o_conc_rhs_ebs_iso(G, p, S->phid, S->cl, S->cs, ydot + (long)S->ic * ng);
         } else if (p->ebs_stencil == AMPE_EBS_ORDER4) {
            o_conc_rhs_ebs_4th(G, p, S->cl, S->cs, S->Dc, S->Dphi, ydot + (long)S->ic * ng);
         } else {
            o_conc_rhs_ebs(G, p, S->cl, S->cs, S->Dc, S->Dphi, ydot + (long)S->ic * ng);
         }
      } else {
         if (recompute) o_conc_coeffs(G, p, phi, S->cl, S->cs, S->Dc, S->Dphi);
         o_conc_rhs(G, p, c, phi, S->Dc, S->Dphi, ydot + (long)S->ic * ng);
      }
   }
   return 0;
}

/* CVSpgmrPrecondSet (source/QuatIntegrator.cc:3300-3376) */
static int psetup(OSim* S, double* y, double gamma)
{
   if (S->Hh) return psetup_hier(S, y, gamma);
   const OGeom* G = &S->G;
   const ampe_model_params* p = &S->p;
   long ng = G->ncg;
   S->st.npe++;
   o_ghost_fill(G, y, S->nc);
   const double* phi = y;
   const double* q = (S->iq >= 0) ? y + (long)S->iq * ng : NULL;
   const double* c = (S->ic >= 0) ? y + (long)S->ic * ng : NULL;
   if (q) o_quat_face_norm2(G, p, q, S->n2);
   if (S->tern) {
      o_kks_tern3(G, p, S->nord, phi, c, S->ph6);
      S->st.kks_calls++;
      S->st.kks_cells += ng;
   } else if (c) {
      int it = o_kks_solve(G, p, phi, c, S->clr, S->csr, S->cl, S->cs);
      S->st.kks_calls++;
      S->st.kks_cells += ng;
      if (it < 0) return -1;
      S->st.kks_iters += it;
   }
   /* phase block: M = M_phi, C = 1 + gamma M w g'', D = -gamma eps^2
    * (source/PhaseFACOps.cc:33-51); with several orders the well scale
    * passed is 0 (source/QuatIntegrator.cc:3321-3329), so C = 1 */
   int kim = p->phi_mobility_type == AMPE_PHI_MOB_KIM;
   if (kim) o_phase_mobility_field(G, p, S->cl, S->cs, S->Mphi);
   {
      ampe_model_params pw = *p;
      if (S->nord > 1) pw.phi_well_scale = 0.0;
      o_phase_fac_C(G, &pw, gamma, phi, kim ? S->Mphi : NULL, S->Cphi);
   }
   /* composition block: D = -gamma D_c (lagged), C = 1, M = M_c
    * (source/ConcFACOps.cc:28-61) */
   if (c)
      for (long f = 0; f < S->nf; f++) {
         /* EBS: D_precond = 0 + D_L + D_A (EBSCompositionRHSStrategy::
          * setDiffusionCoeffForPreconditionerOnPatch) */
         double d = (p->conc_rhs_form == AMPE_CONC_RHS_EBS) ? (0.0 + S->Dc[f]) + S->Dphi[f]
                                                              : S->Dc[f];
         if (S->tern) d = d + S->Db[f];
         S->Dconc[f] = -gamma * d;
      }
   /* orientation block: face coefficients and sqrt(M_q) with ghosts
    * (source/QuatFACOps.cc:735-829) */
   if (q) {
      o_quat_face_coef(G, p, phi, S->n2, S->fc);
      o_quat_mobility(G, p, phi, S->mq);
      for (long n = 0; n < ng; n++) S->sqrtm[n] = sqrt(S->mq[n]);
   }
   S->gamma_setup = gamma;
   return 0;
}

/* Single-level FAC solve of A u = f with the redblack coarse solver
 * (FACPreconditioner::solveSystem, source/samrai/FACPreconditioner.cc:267-429).
 * Returns 0 when the RMS residual is below tol. */
static int fac_solve(OSim* S, const ampe_elliptic_op* op, const double* f,
                     double* u, double tol)
{
   const OGeom* G = &S->G;
   long ng = G->ncg;
   memset(u, 0, sizeof(double) * ng);
   double norm = o_fac_residual(G, op, f, u, S->fac_r, S->fac_flux);
   int iters = 0;
   do {
      memset(S->fac_e, 0, sizeof(double) * ng);
      o_fac_smooth(G, op, S->fac_r, S->fac_e, S->s.fac_sweeps, S->s.fac_tol,
                   S->fac_flux);
      for (long n = 0; n < ng; n++) u[n] = u[n] + S->fac_e[n];
      norm = o_fac_residual(G, op, f, u, S->fac_r, S->fac_flux);
      iters++;
   } while (norm > tol && iters < S->s.fac_max_cycles);
   return norm < tol ? 0 : 1;
}

/* Orientation block operator A = I + gamma sqrt(M) L sqrt(M),
 * L: off-diagonal fc/h^2, diagonal -sum fc/h^2 (set_j_ij2d/3d and
 * set_stencil2d/3d, source/fortran/2d/quatlevelsolver.m4:9-120). */
#define QFX(i, j, k) ((long)(i) + (long)(G->nx + 1) * ((long)(j) + (long)G->ny * (k)))
#define QFY(i, j, k) (G->nfx + (long)(i) + (long)G->nx * ((long)(j) + (long)(G->ny + 1) * (k)))
#define QFZ(i, j, k) (G->nfx + G->nfy + (long)(i) + (long)G->nx * ((long)(j) + (long)G->ny * (k)))

static double quat_apply_row(const OSim* S, double gamma, const double* z,
                             int i, int j, int k, double* center)
{
   const OGeom* G = &S->G;
   const double* fc = S->fc;
   const double* sm = S->sqrtm;
   long c0 = CIG(G, i, j, k);
   double hx = 1.0 / (G->dx[0] * G->dx[0]), hy = 1.0 / (G->dx[1] * G->dx[1]);
   double ow = fc[QFX(i, j, k)] * hx, oe = fc[QFX(i + 1, j, k)] * hx;
   double os = fc[QFY(i, j, k)] * hy, on = fc[QFY(i, j + 1, k)] * hy;
   double diag = -ow - oe - os - on;
   double ob = 0.0, ot = 0.0;
   if (G->ndim == 3) {
      double hz = 1.0 / (G->dx[2] * G->dx[2]);
      ob = fc[QFZ(i, j, k)] * hz;
      ot = fc[QFZ(i, j, k + 1)] * hz;
      diag = diag - ob - ot;
   }
   double fac = gamma * sm[c0];
   *center = fac * diag * sm[c0] + 1.0;
   double r = *center * z[c0] + fac * ow * sm[c0 - 1] * z[c0 - 1] +
              fac * oe * sm[c0 + 1] * z[c0 + 1] +
              fac * os * sm[c0 - G->sy] * z[c0 - G->sy] +
              fac * on * sm[c0 + G->sy] * z[c0 + G->sy];
   if (G->ndim == 3)
      r = r + fac * ob * sm[c0 - G->sz] * z[c0 - G->sz] +
          fac * ot * sm[c0 + G->sz] * z[c0 + G->sz];
   return r;
}

static double quat_rbgs_color(OSim* S, double gamma, const double* rhs,
                              double* z, int color)
{
   const OGeom* G = &S->G;
   double maxres = 0.0;
#pragma omp parallel for reduction(max : maxres)
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++) {
         int x = (G->ndim == 3) ? (j + k) : j;
         int io = ((x - (x / 2) * 2) != color) ? 1 : 0;
         for (int i = io; i < G->nx; i += 2) {
            long c0 = CIG(G, i, j, k);
            double center;
            double az = quat_apply_row(S, gamma, z, i, j, k, &center);
            double res = rhs[c0] - az;
            z[c0] = z[c0] + res / center;
            if (maxres < fabs(res)) maxres = fabs(res);
         }
      }
   return maxres;
}

static double quat_residual(OSim* S, double gamma, const double* rhs,
                            double* z, double* res)
{
   const OGeom* G = &S->G;
   o_ghost_fill(G, z, 1);
   double s = 0.0;
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CIG(G, i, j, k);
            double center;
            double r = rhs[c0] - quat_apply_row(S, gamma, z, i, j, k, &center);
            res[c0] = r;
            s += r * r;
         }
   return s;
}

/* QuatPrecondSolve (source/QuatIntegrator.cc:3594-3650) with
 * QuatSysSolver::solveSystem (source/QuatSysSolver.cc:296-349):
 * divide by sqrt(M), solve, multiply by sqrt(M). */
static int quat_solve(OSim* S, const double* r, double* z, double tol)
{
   const OGeom* G = &S->G;
   long ng = G->ncg, N = (long)G->nx * G->ny * G->nz;
   int ql = S->p.qlen;
   double gamma = S->gamma_setup;
   double* f = S->fac_u;  /* scaled rhs, ql blocks */
   double* res = S->fac_r;
   double* e = S->fac_e;
   for (int m = 0; m < ql; m++)
      for (long n = 0; n < ng; n++)
         f[(long)m * ng + n] = r[(long)m * ng + n] / S->sqrtm[n];
   memset(z, 0, sizeof(double) * ng * ql);
   double s = 0.0;
   for (int m = 0; m < ql; m++)
      s += quat_residual(S, gamma, f + (long)m * ng, z + (long)m * ng,
                         res + (long)m * ng);
   double norm = sqrt(s / ((double)ql * N));
   int iters = 0;
   do {
      memset(e, 0, sizeof(double) * ng * ql);
      for (int m = 0; m < ql; m++) {
         double* em = e + (long)m * ng;
         const double* rm = res + (long)m * ng;
         o_ghost_fill(G, em, 1);
         for (int sw = 0; sw < S->s.quat_sweeps; sw++) {
            o_ghost_fill(G, em, 1);
            quat_rbgs_color(S, gamma, rm, em, 0);
            o_ghost_fill(G, em, 1);
            quat_rbgs_color(S, gamma, rm, em, 1);
            o_ghost_fill(G, em, 1);
         }
      }
      for (long n = 0; n < ng * ql; n++) z[n] = z[n] + e[n];
      s = 0.0;
      for (int m = 0; m < ql; m++)
         s += quat_residual(S, gamma, f + (long)m * ng, z + (long)m * ng,
                            res + (long)m * ng);
      norm = sqrt(s / ((double)ql * N));
      iters++;
   } while (norm > tol && iters < S->s.quat_max_cycles);
   for (int m = 0; m < ql; m++)
      for (long n = 0; n < ng; n++) z[(long)m * ng + n] *= S->sqrtm[n];
   return norm < tol ? 0 : 1;
}

/* CVSpgmrPrecondSolve (source/QuatIntegrator.cc:3666-3771) */
static int psolve(OSim* S, const double* r, double* z, double delta)
{
   if (S->Hh) return psolve_hier(S, r, z, delta);
   const OGeom* G = &S->G;
   const ampe_model_params* p = &S->p;
   long ng = G->ncg;
   int ret = 0;
   S->st.nps++;
   memset(z, 0, sizeof(double) * S->nv);
   ampe_elliptic_op ph = {NULL, -S->gamma_setup * p->epsilon_phi * p->epsilon_phi,
                          S->Cphi, 1.0,
                          p->phi_mobility_type == AMPE_PHI_MOB_KIM ? S->Mphi : NULL,
                          p->phi_mobility,
                          {G->dx[0], G->dx[1], G->dx[2]}};
   /* one phase solve per order (PhasePrecondSolve, :3427-3444) */
   for (int m = 0; m < S->nord; m++)
      if (fac_solve(S, &ph, r + (long)m * ng, z + (long)m * ng, delta)) ret = 1;
   if (S->iq >= 0) {
      if (quat_solve(S, r + (long)S->iq * ng, z + (long)S->iq * ng, delta)) ret = 1;
   }
   /* composition: one block per species with its own diffusion (ConcFACOps
    * depth ic, source/ConcFACOps.cc:38-56; the ternary diagonals are equal) */
   for (int m = 0; m < S->ncs; m++) {
      ampe_elliptic_op co = {S->Dconc, 0.0, NULL, 1.0, NULL, p->conc_mobility,
                             {G->dx[0], G->dx[1], G->dx[2]}};
      if (fac_solve(S, &co, r + (long)(S->ic + m) * ng, z + (long)(S->ic + m) * ng, delta))
         ret = 1;
   }
   return ret;
}

/* SAMRAIVectorReal::dot / weightedRMSNorm with the control volumes of the
 * weight field (level-0 cell volume times the relative weight, zero under a
 * finer level) on a hierarchy; the single-level forms otherwise */
static double hdot(const OSim* S, const double* x, const double* y)
{
   if (!S->Hh) return o_dot(&S->G, S->nc, x, y);
   const OHier* H = S->Hh;
   long ng = o_hier_ncg(H);
   double s = 0.0;
   for (int m = 0; m < S->nc; m++)
      for (long n = 0; n < ng; n++) {
         double w = S->wt[n];
         if (w != 0.0) s += (x[(long)m * ng + n] * y[(long)m * ng + n]) * w;
      }
   const OGeom* G0 = o_hier_geom(H, 0);
   double cv = G0->dx[0] * G0->dx[1] * (G0->ndim == 3 ? G0->dx[2] : 1.0);
   return s * cv;
}

static double wrms(const OSim* S, const double* x)
{
   if (!S->Hh) return o_wrms(&S->G, S->nc, x, S->ewt);
   long ng = o_hier_ncg(S->Hh);
   double s = 0.0;
   for (int m = 0; m < S->nc; m++)
      for (long n = 0; n < ng; n++) {
         double w = S->wt[n];
         if (w == 0.0) continue;
         double v = x[(long)m * ng + n] * S->ewt[(long)m * ng + n];
         s += (v * v) * w;
      }
   return sqrt(s / S->wsum);
}

/* sqrt of the summed control volumes over all components (CVODE's
 * LSNormFactor(-1) norm factor) */
static double norm_factor(const OSim* S)
{
   if (!S->Hh) {
      const OGeom* G = &S->G;
      double cv = G->dx[0] * G->dx[1] * (G->ndim == 3 ? G->dx[2] : 1.0);
      return sqrt((double)S->nc * G->nx * G->ny * G->nz * cv);
   }
   const OGeom* G0 = o_hier_geom(S->Hh, 0);
   double cv = G0->dx[0] * G0->dx[1] * (G0->ndim == 3 ? G0->dx[2] : 1.0);
   return sqrt(S->wsum * cv);
}

/* Jv by difference quotient (CVODE cvLsDQJtimes) and atimes z = v - gamma Jv */
static int atimes(OSim* S, double gamma, double* v, double* out)
{
   double sig = 1.0 / wrms(S, v);
   for (long n = 0; n < S->nv; n++) S->work2[n] = sig * v[n] + S->ycur[n];
   if (rhs_eval(S, S->work2, S->jv, 1)) return -1;
   double siginv = 1.0 / sig;
   for (long n = 0; n < S->nv; n++) {
      double jvn = siginv * S->jv[n] + (-siginv) * S->fcur[n];
      out[n] = v[n] + (-gamma) * jvn;
   }
   return 0;
}

/* SPGMR, left preconditioned, MGS (or classical GS when gstype says so),
 * x0 = 0, s1 = s2 = ewt (SUNLinSol_SPGMR as used by CVODE).  Returns 0 converged, 1 residual
 * reduced but not converged, 2 failure. */
static int spgmr(OSim* S, double gamma, const double* bvec, double* x,
                 double delta, int* nli_out)
{
   int maxl = S->s.maxl;
   long nv = S->nv;
   double** V = S->V;
   double* Hm = S->H;
   double givens[64], g[32], yv[32];
   int ld = maxl + 1;
#define HH(i, j) Hm[(long)(i) * ld + (j)]
   memset(x, 0, sizeof(double) * nv);
   if (psolve(S, bvec, S->work, delta) < 0) return 2;
   for (long n = 0; n < nv; n++) V[0][n] = S->ewt[n] * S->work[n];
   double rho = sqrt(hdot(S, V[0], V[0]));
   double rho0 = rho;
   *nli_out = 0;
   if (rho <= delta) return 0;
   double rinv = 1.0 / rho;
   for (long n = 0; n < nv; n++) V[0][n] *= rinv;
   memset(g, 0, sizeof(g));
   g[0] = rho;
   int l, converged = 0, lused = 0;
   for (l = 0; l < maxl; l++) {
      (*nli_out)++;
      S->st.nli++;
      for (long n = 0; n < nv; n++) S->xk[n] = V[l][n] / S->ewt[n];
      if (atimes(S, gamma, S->xk, V[l + 1])) return 2;
      if (psolve(S, V[l + 1], S->work, delta) < 0) return 2;
      for (long n = 0; n < nv; n++) V[l + 1][n] = S->ewt[n] * S->work[n];
      double hn;
      if (S->s.gstype == AMPE_GS_CLASSICAL) {
         /* classical Gram-Schmidt with one conditional re-orthogonalisation
          * (SUNClassicalGS, SUNDIALS 5.x sundials_iterative.c: all dots of the
          * new vector, one linear combination, re-orthogonalise when the
          * norm dropped by more than a factor 1000) */
         double st[34];
         for (int i = 0; i <= l + 1; i++) st[i] = hdot(S, V[l + 1], V[i]);
         double vk_norm = sqrt(st[l + 1]);
         for (int i = 0; i <= l; i++) HH(i, l) = st[i];
         for (long n = 0; n < nv; n++) {
            double acc = V[l + 1][n];
            for (int i = 0; i <= l; i++) acc = acc + (-st[i]) * V[i][n];
            V[l + 1][n] = acc;
         }
         hn = sqrt(hdot(S, V[l + 1], V[l + 1]));
         if (1000.0 * hn < vk_norm) {
            for (int i = 0; i <= l; i++) st[i] = hdot(S, V[l + 1], V[i]);
            for (int i = 0; i <= l; i++) HH(i, l) += st[i];
            for (long n = 0; n < nv; n++) {
               double acc = V[l + 1][n];
               for (int i = 0; i <= l; i++) acc = acc + (-st[i]) * V[i][n];
               V[l + 1][n] = acc;
            }
            hn = sqrt(hdot(S, V[l + 1], V[l + 1]));
         }
      } else {
         /* modified Gram-Schmidt */
         for (int i = 0; i <= l; i++) {
            double hij = hdot(S, V[i], V[l + 1]);
            HH(i, l) = hij;
            for (long n = 0; n < nv; n++) V[l + 1][n] = V[l + 1][n] + (-hij) * V[i][n];
         }
         hn = sqrt(hdot(S, V[l + 1], V[l + 1]));
      }
      HH(l + 1, l) = hn;
      /* apply previous Givens rotations to column l, then a new one */
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
      } else if (fabs(bb) > fabs(a)) {
         double t = a / bb;
         s = -1.0 / sqrt(1.0 + t * t);
         c = -s * t;
      } else {
         double t = bb / a;
         c = 1.0 / sqrt(1.0 + t * t);
         s = -c * t;
      }
      givens[2 * l] = c;
      givens[2 * l + 1] = s;
      HH(l, l) = c * a - s * bb;
      HH(l + 1, l) = 0.0;
      g[l + 1] = s * g[l];
      g[l] = c * g[l];
      rho = fabs(g[l + 1]);
      lused = l + 1;
      if (dbg()) fprintf(stderr, "  gmres l=%d rho=%.3e (rho0=%.3e delta=%.3e)\n", l, rho, rho0, delta);
      if (rho <= delta) {
         converged = 1;
         break;
      }
      if (hn == 0.0) break;
      double hinv = 1.0 / hn;
      for (long n = 0; n < nv; n++) V[l + 1][n] *= hinv;
   }
   /* back substitution H y = g */
   for (int i = lused - 1; i >= 0; i--) {
      double sacc = g[i];
      for (int j = i + 1; j < lused; j++) sacc -= HH(i, j) * yv[j];
      yv[i] = sacc / HH(i, i);
   }
   for (int i = 0; i < lused; i++)
      for (long n = 0; n < nv; n++) x[n] = x[n] + yv[i] * V[i][n];
   for (long n = 0; n < nv; n++) x[n] = x[n] / S->ewt[n];
#undef HH
   if (converged) return 0;
   return (rho < rho0) ? 1 : 2;
}

/* Newton iteration for one step attempt with step h: returns 0 converged,
 * 1 convergence failure, -1 fatal.  *did_setup reports a fresh lsetup. */
static int newton(OSim* S, double gamma, double delta, double tq4, int force_setup,
                  int* did_setup)
{
   long nv = S->nv;
   int call_setup = force_setup || S->nst == 0 ||
                    fabs(gamma / S->gamma_setup - 1.0) > 0.3 ||
                    S->nst >= S->nstlp + S->s.msbp;
   *did_setup = 0;
   memcpy(S->ycur, S->ypred, sizeof(double) * nv);
   double delp = 0.0;
   for (int m = 0; m < S->s.max_newton; m++) {
      S->st.nni++;
      if (rhs_eval(S, S->ycur, S->fcur, 0)) return -1;
      for (long n = 0; n < nv; n++)
         S->b[n] = gamma * S->fcur[n] - (S->ycur[n] - S->ahist[n]);
      if (m == 0 && call_setup) {
         if (psetup(S, S->ycur, gamma)) return -1;
         S->nstlp = S->nst;
         S->crate = 1.0;
         *did_setup = 1;
      }
      int nli = 0;
      int ls = spgmr(S, gamma, S->b, S->del, delta, &nli);
      if (ls == 2 || (ls == 1 && m > 0)) {
         S->st.nlcf++;
         return 1;
      }
      double del = wrms(S, S->del);
      for (long n = 0; n < nv; n++) S->ycur[n] = S->ycur[n] + S->del[n];
      if (m > 0) {
         double r = del / delp;
         S->crate = (0.3 * S->crate > r) ? 0.3 * S->crate : r;
      }
      double dcon = del * (S->crate < 1.0 ? S->crate : 1.0) / tq4;
      if (dbg()) fprintf(stderr, " newton m=%d ls=%d del=%.3e crate=%.3e\n", m, ls, del, S->crate);
      if (dcon <= 1.0) return 0;
      if (m >= 1 && del > 2.0 * delp) return 1;
      delp = del;
   }
   return 1;
}

/* One accepted BDF step: BDF1 on the first step, variable-step BDF2 after
 * (step ratio w = h/hprev: gamma = h(1+w)/(1+2w),
 * a = ((1+w)^2 y_n - w^2 y_{n-1})/(1+2w), predictor y_n + w (y_n - y_{n-1})).
 * A Newton failure with a fresh preconditioner cuts h by 4 (CVODE's
 * convergence-failure factor); h then regrows by 2 per step up to dt.
 * Returns 0 on success, 1 if h fell below dt*1e-8, -1 fatal. */
static int bdf_step(OSim* S)
{
   const OGeom* G = &S->G;
   long nv = S->nv, ng = G->ncg;
   /* resetRefPhaseConcentrations (source/QuatModel.cc:5218-5235) */
   if (S->ic >= 0) {
      memcpy(S->clr, S->cl, sizeof(double) * ng);
      memcpy(S->csr, S->cs, sizeof(double) * ng);
   }
   /* nrmfac = sqrt(sum of control volumes over all components)
    * (CVodeSetLSNormFactor(-1): sqrt(N_VDotProd(1,1))) */
   double nrmfac = norm_factor(S);
   double delta = S->s.eplifac * S->s.nlscoef * nrmfac;
   int force_setup = 0;
   for (;;) {
      double h = S->h, gamma;
      if (S->tramp && S->dTdt != 0.0) {
         S->p.temperature = ramp_temperature(S, S->t + h);
         if (S->has_tab) o_calphad_at(&S->tab, S->p.temperature, S->p.calphad);
      }
      if (!S->have_prev) {
         /* zeroth-order predictor on the first step: the step is imposed,
          * not chosen by CVODE's initial-step heuristic, so y0 + h f(y0) can
          * be far outside the Newton ball for stiff relaxing modes */
         memcpy(S->ypred, S->y, sizeof(double) * nv);
      } else {
         double w = h / S->hprev;
         for (long n = 0; n < nv; n++)
            S->ypred[n] = S->y[n] + w * (S->y[n] - S->yprev[n]);
      }
      if (!S->have_prev || S->s.max_order < 2) {
         gamma = h;
         memcpy(S->ahist, S->y, sizeof(double) * nv);
      } else {
         double w = h / S->hprev;
         double den = 1.0 + 2.0 * w;
         gamma = h * (1.0 + w) / den;
         double c1 = (1.0 + w) * (1.0 + w) / den, c2 = w * w / den;
         for (long n = 0; n < nv; n++) S->ahist[n] = c1 * S->y[n] - c2 * S->yprev[n];
      }
      int did_setup = 0;
      int r = newton(S, gamma, delta, S->s.nlscoef, force_setup, &did_setup);
      if (r < 0) return -1;
      if (r == 0) {
         for (long n = 0; n < nv; n++) S->acor[n] = S->ycur[n] - S->ypred[n];
         if (S->iq >= 0 && S->p.qlen > 1) {
            memset(S->corr, 0, sizeof(double) * nv);
            project_q(S, S->ycur + (long)S->iq * ng, S->corr + (long)S->iq * ng,
                      S->acor + (long)S->iq * ng);
            for (long n = 0; n < nv; n++) S->ycur[n] = S->ycur[n] + S->corr[n];
         }
         S->st.last_acnrm = wrms(S, S->acor);
         memcpy(S->yprev, S->y, sizeof(double) * nv);
         memcpy(S->y, S->ycur, sizeof(double) * nv);
         S->have_prev = 1;
         S->hprev = h;
         S->t += h;
         S->nst++;
         S->st.nst++;
         S->h = (2.0 * h < S->s.dt) ? 2.0 * h : S->s.dt;
         update_ewt(S, S->y);
         return 0;
      }
      S->st.ncfn++;
      if (!did_setup) {
         force_setup = 1;
         continue;
      }
      S->h = 0.25 * h;
      force_setup = 1;
      if (S->h < S->s.dt * 1.0e-8) return 1;
   }
}

/* ------------------------------------------------------------------------
 * Adaptive BDF (s.adaptive = 1): SUNDIALS CVODE 5.x as QuatIntegrator
 * configures it (source/QuatIntegrator.cc:1571-1596), restated from CVODE's
 * published algorithm (Hindmarsh et al., ACM TOMS 31 (2005) 363; CVODE user
 * guide, "Mathematical considerations") for orders 1..2:
 *   - Nordsieck history z = (y, h y', h^2 y''/2), predicted by the Pascal
 *     matrix (cvPredict) and restored by its inverse after a failure;
 *   - fixed-leading-coefficient BDF coefficients l, error constants tq
 *     (cvSetBDF / cvSetTqBDF) from the step history tau;
 *   - corrector y_n = y_pred + acor, acor = gamma f(y_n) - z1 / l1,
 *     gamma = h / l1; Newton test del min(1, crate) <= tq4 = nlscoef / tq2,
 *     linear tolerance eplifac tq4 sqrt(sum cv);
 *   - projection of q after convergence, the error estimate projected with it;
 *   - error test dsm = ||acor||_wrms tq2 <= 1; failure: eta = max(0.1,
 *     1/((6 dsm)^(1/L) + 1e-6)), at most 0.2 from the second failure, order
 *     reduction (or reload of z1 = h f at order 1) after the third, give up
 *     after 7; convergence failure: eta = 1/4, give up after 10;
 *   - next step (cvPrepareNextStep): eta_q, and every q + 1 steps eta_{q-1}
 *     (6 ||z_q|| tq1) and eta_{q+1} (10 ||acor - c z_saved|| tq3), the largest
 *     taken when >= 1.5, bounded by etamax (1e4 on the first step, 10 after,
 *     1 after a failed attempt) and h <= dt.
 * Parity with SUNDIALS itself is unpinned (the library is absent); the device
 * integrator implements the same algorithm.
 * ---------------------------------------------------------------------- */
static void axpby(long n, double a, const double* x, double b, const double* y, double* z)
{
   for (long i = 0; i < n; i++) z[i] = a * x[i] + b * y[i];
}

static double* zcol(OSim* S, int j) { return j == 0 ? S->y : (j == 1 ? S->z1 : S->z2); }

static void bdf_coeffs(OSim* S, double h)
{
   int q = S->q;
   double l0 = 1.0, l1 = 1.0, l2 = 0.0;
   double alpha0 = -1.0, alpha0_hat = -1.0, xi_inv = 1.0, xistar_inv = 1.0;
   double hsum = h;
   if (q == 2) {
      alpha0 -= 1.0 / 2.0;
      xistar_inv = -l1 - alpha0;
      hsum += S->tau[1];
      xi_inv = h / hsum;
      alpha0_hat = -l1 - xi_inv;
      l2 += l1 * xistar_inv;
      l1 += l0 * xistar_inv;
   }
   S->l[0] = l0;
   S->l[1] = l1;
   S->l[2] = l2;
   double lq = (q == 2) ? l2 : l1;
   double A1 = 1.0 - alpha0_hat + alpha0;
   double A2 = 1.0 + q * A1;
   S->tq[2] = fabs(A1 / (alpha0 * A2));
   S->tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
   if (S->qwait == 1) {
      if (q > 1) {
         double Cc = xistar_inv / lq;
         double A3 = alpha0 + 1.0 / q;
         double A4 = alpha0_hat + xi_inv;
         double Cpinv = (1.0 - A4 + A3) / A3;
         S->tq[1] = fabs(Cc * Cpinv);
      } else {
         S->tq[1] = 1.0;
      }
      hsum += S->tau[q];
      xi_inv = h / hsum;
      double A5 = alpha0 - 1.0 / (q + 1);
      double A6 = alpha0_hat - xi_inv;
      double Cppinv = (1.0 - A6 + A5) / A2;
      S->tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
   }
   S->tq[4] = S->s.nlscoef / S->tq[2];
}

static double eta_of(double bias, double err, int p)
{
   return 1.0 / (pow(bias * err, 1.0 / p) + 1.0e-6);
}

static void bdf_predict(OSim* S, double sign)
{
   for (int k = 1; k <= S->q; k++)
      for (int j = S->q; j >= k; j--) axpby(S->nv, 1.0, zcol(S, j - 1), sign, zcol(S, j), zcol(S, j - 1));
}

static void bdf_rescale(OSim* S, double eta)
{
   double factor = eta;
   for (int j = 1; j <= S->q; j++) {
      axpby(S->nv, factor, zcol(S, j), 0.0, zcol(S, j), zcol(S, j));
      factor *= eta;
   }
   S->h = S->hscale * eta;
   S->hscale = S->h;
}

static void bdf_set_eta(OSim* S, double h)
{
   if (S->eta < 1.5) {
      S->eta = 1.0;
      S->hprime = h;
   } else {
      double hmax = S->s.dt;
      S->eta = S->eta < S->etamax ? S->eta : S->etamax;
      double d = fabs(h) * (1.0 / hmax) * S->eta;
      S->eta /= (d > 1.0 ? d : 1.0);
      S->hprime = h * S->eta;
   }
}

static void bdf_prepare_next(OSim* S, double dsm, const double* acor)
{
   double h = S->h;
   int qmax = S->s.max_order;
   if (S->etamax == 1.0) {
      S->qwait = S->qwait > 2 ? S->qwait : 2;
      S->qprime = S->q;
      S->hprime = h;
      S->eta = 1.0;
      return;
   }
   double etaq = eta_of(6.0, dsm, S->L);
   if (S->qwait != 0) {
      S->eta = etaq;
      S->qprime = S->q;
      bdf_set_eta(S, h);
      return;
   }
   S->qwait = 2;
   double etaqm1 = 0.0, etaqp1 = 0.0;
   if (S->q > 1) etaqm1 = eta_of(6.0, wrms(S, zcol(S, S->q)) * S->tq[1], S->q);
   if (S->q != qmax && S->saved_tq5 != 0.0) {
      double ratio = h / S->tau[2], pw = 1.0;
      for (int i = 0; i < S->L; i++) pw *= ratio;
      double cquot = (S->tq[5] / S->saved_tq5) * pw;
      axpby(S->nv, -cquot, S->z2, 1.0, acor, S->work2);
      etaqp1 = eta_of(10.0, wrms(S, S->work2) * S->tq[3], S->L + 1);
   }
   double etam = etaq > etaqp1 ? etaq : etaqp1;
   etam = etaqm1 > etam ? etaqm1 : etam;
   if (etam < 1.5) {
      S->eta = 1.0;
      S->qprime = S->q;
   } else if (etam == etaq) {
      S->eta = etaq;
      S->qprime = S->q;
   } else if (etam == etaqm1) {
      S->eta = etaqm1;
      S->qprime = S->q - 1;
   } else {
      S->eta = etaqp1;
      S->qprime = S->q + 1;
      memcpy(S->z2, acor, sizeof(double) * S->nv);
   }
   bdf_set_eta(S, h);
}

static void bdf_adjust_order(OSim* S, int qnew)
{
   if (qnew > S->q) memset(S->z2, 0, sizeof(double) * S->nv);
   S->q = qnew;
   S->L = qnew + 1;
   S->qwait = S->L;
}

static int bdf_load_z1(OSim* S, double t)
{
   if (S->tramp && S->dTdt != 0.0) {
      S->p.temperature = ramp_temperature(S, t);
      if (S->has_tab) o_calphad_at(&S->tab, S->p.temperature, S->p.calphad);
   }
   if (rhs_eval(S, S->y, S->fcur, 0)) return -1;
   axpby(S->nv, S->h, S->fcur, 0.0, S->fcur, S->z1);
   return 0;
}

static int bdf_adaptive_step(OSim* S)
{
   const OGeom* G = &S->G;
   long nv = S->nv, ng = G->ncg;
   int qmax = S->s.max_order;
   double hmin = S->s.dt * 1.0e-8;
   if (S->ic >= 0) {
      memcpy(S->clr, S->cl, sizeof(double) * ng);
      memcpy(S->csr, S->cs, sizeof(double) * ng);
   }
   if (!S->started) {
      S->q = S->qprime = 1;
      S->L = 2;
      S->qwait = 2;
      for (int i = 0; i < 4; i++) S->tau[i] = 0.0;
      S->hscale = S->hprime = S->h;
      S->eta = 1.0;
      S->etamax = 1.0e4;
      S->saved_tq5 = 0.0;
      if (bdf_load_z1(S, S->t)) return -1;
      S->started = 1;
   } else if (S->hprime != S->h) {
      if (S->qprime != S->q) bdf_adjust_order(S, S->qprime);
      bdf_rescale(S, S->eta);
   }
   double nrmfac = norm_factor(S);
   int nef = 0, ncf = 0, force_setup = 0, predicted = 0;
   double gamma = 0.0, delta = 0.0;
   for (;;) {
      double h = S->h;
      if (!predicted) {
         if (S->tramp && S->dTdt != 0.0) {
            S->p.temperature = ramp_temperature(S, S->t + h);
            if (S->has_tab) o_calphad_at(&S->tab, S->p.temperature, S->p.calphad);
         }
         bdf_predict(S, 1.0);
         memcpy(S->ypred, S->y, sizeof(double) * nv);
         bdf_coeffs(S, h);
         gamma = h / S->l[1];
         axpby(nv, 1.0, S->ypred, -1.0 / S->l[1], S->z1, S->ahist);
         delta = S->s.eplifac * S->tq[4] * nrmfac;
         predicted = 1;
      }
      int did_setup = 0;
      int r = newton(S, gamma, delta, S->tq[4], force_setup, &did_setup);
      if (r < 0) return -1;
      if (r != 0) {
         S->st.ncfn++;
         if (!did_setup) {
            force_setup = 1;
            continue;
         }
         bdf_predict(S, -1.0);
         predicted = 0;
         ncf++;
         S->etamax = 1.0;
         if (ncf >= 10 || h <= hmin) return 1;
         bdf_rescale(S, 0.25);
         force_setup = 1;
         continue;
      }
      axpby(nv, 1.0, S->ycur, -1.0, S->ypred, S->acor);
      if (S->iq >= 0 && S->p.qlen > 1) {
         memset(S->corr, 0, sizeof(double) * nv);
         project_q(S, S->ycur + (long)S->iq * ng, S->corr + (long)S->iq * ng,
                   S->acor + (long)S->iq * ng);
         axpby(nv, 1.0, S->ycur, 1.0, S->corr, S->ycur);
      }
      double acnrm = wrms(S, S->acor);
      S->st.last_acnrm = acnrm;
      double dsm = acnrm * S->tq[2];
      if (dbg())
         fprintf(stderr, "adaptive t=%.6e h=%.6e q=%d dsm=%.4e\n", S->t, h, S->q, dsm);
      if (dsm > 1.0) {
         bdf_predict(S, -1.0);
         predicted = 0;
         nef++;
         S->st.netf++;
         S->etamax = 1.0;
         if (nef >= 7 || h <= hmin) return 1;
         if (nef <= 3) {
            double eta = eta_of(6.0, dsm, S->L);
            eta = eta > 0.1 ? eta : 0.1;
            if (nef >= 2 && eta > 0.2) eta = 0.2;
            bdf_rescale(S, eta);
         } else if (S->q > 1) {
            bdf_adjust_order(S, S->q - 1);
            bdf_rescale(S, 0.1);
         } else {
            S->h = 0.1 * h;
            S->hscale = S->h;
            S->qwait = 10;
            if (bdf_load_z1(S, S->t)) return -1;
         }
         continue;
      }
      axpby(nv, 1.0, S->ycur, -1.0, S->ypred, S->work);
      S->nst++;
      S->st.nst++;
      for (int i = S->q; i >= 2; i--) S->tau[i] = S->tau[i - 1];
      if (S->q == 1 && S->nst > 1) S->tau[2] = S->tau[1];
      S->tau[1] = h;
      for (int j = 0; j <= S->q; j++) axpby(nv, S->l[j], S->work, 1.0, zcol(S, j), zcol(S, j));
      S->qwait--;
      if (S->qwait == 1 && S->q != qmax) {
         memcpy(S->z2, S->work, sizeof(double) * nv);
         S->saved_tq5 = S->tq[5];
      }
      S->t += h;
      S->st.q = S->q;
      bdf_prepare_next(S, dsm, S->work);
      S->etamax = 10.0;
      update_ewt(S, S->y);
      return 0;
   }
}


/* ------------------------------------------------------------------------
 * The integrator on a patch hierarchy: composite state vectors, the
 * callbacks of QuatIntegrator over all levels (setCoefficients ->
 * coarsenData of the iterate, fillScratch with LINEAR_REFINE coarse-fine
 * ghosts, source/QuatIntegrator.cc:2994-3083, 2873-2957), composite
 * right-hand sides, FAC preconditioner blocks over levels, and the
 * control-volume weighted dot products and norms.
 * ---------------------------------------------------------------------- */
static void hier_sync_fill(OSim* S, double* y)
{
   const OHier* H = S->Hh;
   long ng = o_hier_ncg(H);
   o_hier_coarsen(H, y, S->nord, 0);
   if (S->ic >= 0) o_hier_coarsen(H, y + (long)S->ic * ng, S->ncs, 0);
   if (S->iq >= 0) o_hier_coarsen(H, y + (long)S->iq * ng, S->p.qlen, 1);
   o_hier_ghost_fill(H, y, S->nc, 2);
}

static int hier_kks(OSim* S, const double* y)
{
   const OHier* H = S->Hh;
   long ng = o_hier_ncg(H);
   if (S->tern) {
      /* closed form, cell by cell over the composite (ghosts included) */
      OGeom Gc = *o_hier_geom(H, 0);
      Gc.ncg = ng;
      o_kks_tern3(&Gc, &S->p, S->nord, y, y + (long)S->ic * ng, S->ph6);
      S->st.kks_calls++;
      S->st.kks_cells += ng;
      return 0;
   }
   int np = o_hier_npatch(H);
   long* coff = (long*)malloc(sizeof(long) * np);
   o_hier_offsets(H, coff, NULL);
   const double* c = y + (long)S->ic * ng;
   long tot = 0;
   int fail = 0;
   for (int p = 0; p < np; p++) {
      const OGeom* G = o_hier_geom(H, p);
      long o = coff[p];
      int it = o_kks_solve(G, &S->p, y + o, c + o, S->clr + o, S->csr + o, S->cl + o, S->cs + o);
      if (it < 0) fail = 1;
      else tot += it;
   }
   free(coff);
   S->st.kks_calls++;
   S->st.kks_cells += ng;
   if (fail) return -1;
   S->st.kks_iters += tot;
   return 0;
}

static int rhs_eval_hier(OSim* S, double* y, double* ydot, int fd)
{
   const OHier* H = S->Hh;
   const ampe_model_params* p = &S->p;
   long ng = o_hier_ncg(H);
   int recompute = (!fd) || (!S->s.lag_quat_sidegrad);
   S->st.nfe++;
   hier_sync_fill(S, y);
   const double* phi = y;
   const double* q = (S->iq >= 0) ? y + (long)S->iq * ng : NULL;
   const double* c = (S->ic >= 0) ? y + (long)S->ic * ng : NULL;
   if (S->tern) {
      if (hier_kks(S, y)) return -1;
      o_hier_phase_rhs_tern3(H, p, S->nord, phi, S->ph6, ydot);
      if (recompute) o_hier_ebs3_coeffs(H, p, S->nord, phi, S->Dc, S->Dphi, S->Db);
      o_hier_conc_rhs_ebs3(H, p, S->ph6, S->Dc, S->Dphi, S->Db, ydot + (long)S->ic * ng);
      return 0;
   }
   if (S->nord > 1) {
      o_hier_phase_rhs_multiorder(H, p, S->nord, phi, ydot);
      return 0;
   }
   if (q && recompute) o_hier_quat_face_norm2(H, p, q, S->n2);
   if (c && hier_kks(S, y)) return -1;
   o_hier_phase_rhs(H, p, phi, q, S->cl, S->cs, ydot);
   if (q) o_hier_quat_rhs(H, p, phi, q, S->n2, ydot + (long)S->iq * ng);
   if (c) {
      if (recompute) o_hier_conc_coeffs(H, p, phi, S->cl, S->cs, S->Dc, S->Dphi);
      o_hier_conc_rhs(H, p, c, phi, S->Dc, S->Dphi, ydot + (long)S->ic * ng);
   }
   return 0;
}

static int psetup_hier(OSim* S, double* y, double gamma)
{
   const OHier* H = S->Hh;
   const ampe_model_params* p0 = &S->p;
   long ng = o_hier_ncg(H);
   int np = o_hier_npatch(H);
   S->st.npe++;
   hier_sync_fill(S, y);
   const double* phi = y;
   const double* q = (S->iq >= 0) ? y + (long)S->iq * ng : NULL;
   const double* c = (S->ic >= 0) ? y + (long)S->ic * ng : NULL;
   if (q) o_hier_quat_face_norm2(H, p0, q, S->n2);
   if (c && hier_kks(S, y)) return -1;
   long* coff = (long*)malloc(sizeof(long) * np);
   long* foff = (long*)malloc(sizeof(long) * np);
   o_hier_offsets(H, coff, foff);
   int kim = p0->phi_mobility_type == AMPE_PHI_MOB_KIM;
   for (int pp = 0; pp < np; pp++) {
      const OGeom* G = o_hier_geom(H, pp);
      ampe_model_params p = *p0;
      for (int a = 0; a < 3; a++) p.dx[a] = G->dx[a];
      long o = coff[pp];
      /* several orders: the well scale passed is 0 (C = 1) */
      if (S->nord > 1) p.phi_well_scale = 0.0;
      if (kim) o_phase_mobility_field(G, &p, S->cl + o, S->cs + o, S->Mphi + o);
      o_phase_fac_C(G, &p, gamma, phi + o, kim ? S->Mphi + o : NULL, S->Cphi + o);
      if (q) {
         o_quat_face_coef(G, &p, phi + o, S->n2 + foff[pp], S->fc + foff[pp]);
         o_quat_mobility(G, &p, phi + o, S->mq + o);
      }
   }
   free(coff);
   free(foff);
   if (c)
      for (long f = 0; f < S->nf; f++) {
         double d = S->tern ? ((0.0 + S->Dc[f]) + S->Dphi[f]) + S->Db[f] : S->Dc[f];
         S->Dconc[f] = -gamma * d;
      }
   if (q)
      for (long n = 0; n < ng; n++) S->sqrtm[n] = sqrt(S->mq[n]);
   S->gamma_setup = gamma;
   return 0;
}

/* CVSpgmrPrecondSolve with the FAC blocks cycled over the levels
 * (PhaseFACSolver / ConcFACSolver: EllipticFACOps with Ewing's fix;
 * QuatSysSolver: divide by sqrt(M), QuatFACOps, multiply) */
static int psolve_hier(OSim* S, const double* r, double* z, double delta)
{
   const OHier* H = S->Hh;
   const ampe_model_params* p = &S->p;
   long ng = o_hier_ncg(H);
   int ret = 0;
   S->st.nps++;
   memset(z, 0, sizeof(double) * S->nv);
   OHFacParams fp = {S->s.fac_max_cycles, 1, 1, S->s.fac_sweeps, S->s.fac_tol};
   OHBlock ph = {0, NULL, -S->gamma_setup * p->epsilon_phi * p->epsilon_phi, S->Cphi, 1.0,
                 p->phi_mobility_type == AMPE_PHI_MOB_KIM ? S->Mphi : NULL, p->phi_mobility,
                 NULL, 0.0, 1};
   /* one phase block per order (PhasePrecondSolve, source/QuatIntegrator.cc:3427-3444) */
   for (int m = 0; m < S->nord; m++)
      if (o_hier_fac_solve(H, &ph, &fp, r + (long)m * ng, z + (long)m * ng, 1, delta, NULL,
                           NULL))
         ret = 1;
   if (S->iq >= 0) {
      int ql = p->qlen;
      double* f = S->fac_u;
      for (int m = 0; m < ql; m++)
         for (long n = 0; n < ng; n++)
            f[(long)m * ng + n] = r[(long)(S->iq + m) * ng + n] / S->sqrtm[n];
      double* u = z + (long)S->iq * ng;
      OHFacParams qp = {S->s.quat_max_cycles, 1, 1, S->s.quat_sweeps, -1.0};
      OHBlock qb = {1, S->fc, 0.0, NULL, 1.0, NULL, 1.0, S->sqrtm, S->gamma_setup, 0};
      if (o_hier_fac_solve(H, &qb, &qp, f, u, ql, delta, NULL, NULL)) ret = 1;
      for (int m = 0; m < ql; m++)
         for (long n = 0; n < ng; n++) u[(long)m * ng + n] *= S->sqrtm[n];
   }
   /* one composition block per species */
   for (int m = 0; m < S->ncs; m++) {
      OHBlock co = {0, S->Dconc, 0.0, NULL, 1.0, NULL, p->conc_mobility, NULL, 0.0, 1};
      if (o_hier_fac_solve(H, &co, &fp, r + (long)(S->ic + m) * ng, z + (long)(S->ic + m) * ng,
                           1, delta, NULL, NULL))
         ret = 1;
   }
   return ret;
}

/* project2d/3d on every interior cell (all patches) */
static void project_q(OSim* S, const double* q, double* corr, double* err)
{
   if (!S->Hh) {
      o_quat_project(&S->G, &S->p, q, corr, err);
      return;
   }
   const OHier* H = S->Hh;
   long ng = o_hier_ncg(H);
   int ql = S->p.qlen;
   for (long n = 0; n < ng; n++) {
      if (!S->interior[n]) continue;
      double qv[4], cv[4], fac = 0.0;
      for (int m = 0; m < ql; m++) {
         qv[m] = q[(long)m * ng + n];
         fac = fac + qv[m] * qv[m];
      }
      fac = 1.0 / sqrt(fac);
      for (int m = 0; m < ql; m++) cv[m] = qv[m] * fac;
      fac = 0.0;
      for (int m = 0; m < ql; m++) fac = fac + cv[m] * err[(long)m * ng + n];
      for (int m = 0; m < ql; m++) err[(long)m * ng + n] = err[(long)m * ng + n] - cv[m] * fac;
      for (int m = 0; m < ql; m++) corr[(long)m * ng + n] = cv[m] - qv[m];
   }
}

OSim* o_sim_create_hier(const ampe_model_params* p, const ampe_solver_params* s, OHier* H)
{
   const int tern = p->fe_model == AMPE_FE_QUADRATIC_TERNARY_3PH;
   if (p->with_conc && !tern && p->conc_rhs_form != AMPE_CONC_RHS_KKS) return NULL;
   if (p->evolve_quat && p->qlen > 4) return NULL;
   OSim* S = o_sim_create(p, s, 1, 1, 1);
   if (!S) return NULL;
   /* resize every field to the composite */
   long ng = o_hier_ncg(H), nf = o_hier_nf(H);
   S->Hh = H;
   S->G.ncg = ng;
   S->nv = (long)S->nc * ng;
   S->nf = nf;
   double** vs[] = {&S->y, &S->yprev, &S->ypred, &S->ycur, &S->ahist, &S->fcur,
                    &S->b, &S->del, &S->ewt, &S->work, &S->work2, &S->jv,
                    &S->xk, &S->corr, &S->acor, &S->z1, &S->z2};
   for (size_t i = 0; i < sizeof(vs) / sizeof(vs[0]); i++) {
      free(*vs[i]);
      *vs[i] = vec(S->nv);
   }
   for (int i = 0; i <= s->maxl; i++) {
      free(S->V[i]);
      S->V[i] = vec(S->nv);
   }
   double** cs[] = {&S->cl, &S->cs, &S->clr, &S->csr, &S->Cphi, &S->Mphi, &S->phid,
                    &S->sqrtm, &S->mq};
   for (size_t i = 0; i < sizeof(cs) / sizeof(cs[0]); i++) {
      free(*cs[i]);
      *cs[i] = vec(ng);
   }
   int ql = p->evolve_quat ? p->qlen : 1;
   free(S->fac_u);
   free(S->fac_e);
   free(S->fac_r);
   S->fac_u = vec(ng * ql);
   S->fac_e = vec(ng * ql);
   S->fac_r = vec(ng * ql);
   free(S->ph6);
   S->ph6 = vec(6 * ng);
   double** fs[] = {&S->n2, &S->Dc, &S->Dphi, &S->Dconc, &S->fc, &S->fac_flux, &S->Db};
   for (size_t i = 0; i < sizeof(fs) / sizeof(fs[0]); i++) {
      free(*fs[i]);
      *fs[i] = vec(nf);
   }
   S->wt = vec(ng);
   o_hier_weights(H, S->wt);
   double sw = 0.0;
   for (long n = 0; n < ng; n++) sw += S->wt[n];
   S->wsum = (double)S->nc * sw;
   /* interior mask of the composite */
   S->interior = (unsigned char*)calloc((size_t)ng, 1);
   int np = o_hier_npatch(H);
   long* coff = (long*)malloc(sizeof(long) * np);
   o_hier_offsets(H, coff, NULL);
   for (int pp = 0; pp < np; pp++) {
      const OGeom* G = o_hier_geom(H, pp);
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) S->interior[coff[pp] + CIG(G, i, j, k)] = 1;
   }
   free(coff);
   return S;
}

/* composite state (nc blocks of the composite cell layout; ghosts ignored):
 * coarsened, ghosts filled, initial phase concentrations from cL = cS = c */
void o_sim_set_state_composite(OSim* S, const double* y)
{
   memcpy(S->y, y, sizeof(double) * S->nv);
   hier_sync_fill(S, S->y);
   if (S->tern) {
      hier_kks(S, S->y);
   } else if (S->ic >= 0) {
      long ng = o_hier_ncg(S->Hh);
      const double* cc = S->y + (long)S->ic * ng;
      memcpy(S->clr, cc, sizeof(double) * ng);
      memcpy(S->csr, cc, sizeof(double) * ng);
      hier_kks(S, S->y);
   }
   S->t = 0.0;
   S->nst = 0;
   S->have_prev = 0;
   S->h = first_step(&S->s);
   S->started = 0;
   memset(&S->st, 0, sizeof(S->st));
   update_ewt(S, S->y);
}

/* the integrator restarted at time t after a set_state (a regrid:
 * QuatIntegrator::resetIntegrator, source/QuatIntegrator.cc:1676-1730, with
 * setSundialsOptions' initial time and step, :1571-1580); h0 > 0 is the first
 * step, else the solver's own first step */
void o_sim_set_time(OSim* S, double t, double h0)
{
   S->t = t;
   if (h0 > 0.0) S->h = h0;
}

/* the state with every finer level coarsened onto the cells it covers (the
 * output QuatIntegrator coarsens after each step, :2190-2196) */
void o_sim_get_state_composite(OSim* S, double* y)
{
   memcpy(y, S->y, sizeof(double) * S->nv);
   long ng = o_hier_ncg(S->Hh);
   o_hier_coarsen(S->Hh, y, S->nord, 0);
   if (S->ic >= 0) o_hier_coarsen(S->Hh, y + (long)S->ic * ng, S->ncs, 0);
   if (S->iq >= 0) o_hier_coarsen(S->Hh, y + (long)S->iq * ng, S->p.qlen, 1);
}

int o_sim_advance(OSim* S, int nsteps, ampe_stats* st)
{
   int ret = 0;
   for (int n = 0; n < nsteps; n++) {
      ret = S->s.adaptive ? bdf_adaptive_step(S) : bdf_step(S);
      if (ret) break;
   }
   S->st.t = S->t;
   S->st.h = S->s.adaptive ? S->hprime : S->h;
   if (!S->s.adaptive) S->st.q = (S->have_prev && S->s.max_order >= 2) ? 2 : 1;
   if (st) *st = S->st;
   return ret;
}
