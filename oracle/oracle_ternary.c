/*
 * Three phases (liquid, solid A, solid B) on norderp > 1 order parameters
 * with a ternary quadratic model -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated, each function citing the reference:
 *  - phase concentrations (QuadraticEquilibriumThreePhasesTernaryStrategyMultiOrder::
 *    computePhaseConcentrationsOnPatch, source/QuadraticEquilibriumThreePhases
 *    TernaryStrategyMultiOrder.cc:49-181): phase weights h = (phi_L^2,
 *    sum_A phi^2, sum_B phi^2) / sum phi^2, then the KKS conditions of
 *    Thermo4PFM's QuadraticFreeEnergyFunctionsTernaryThreePhase
 *    (computePhaseConcentrations; Thermo4PFM is not in the reference tree).
 *    For f_a(c) = sum_k A_ak (c_k - e_ak)^2 the conditions -- one diffusion
 *    potential mu_k = 2 A_ak (x_ak - e_ak) in every phase and
 *    c_k = sum_a g(h_a) x_ak (g: conc_interp) -- have the closed form
 *    mu_k = (c_k - sum_a g_a e_ak) / sum_a (g_a / (2 A_ak)); the operation
 *    order of Thermo4PFM's solver is "parity unpinned", the solution is
 *    pinned by the two conditions (tests/test_ternary_oracle.py);
 *  - the free energies, potentials and Moelans' driving force of
 *    QuadraticFreeEnergyStrategyMultiOrderTernaryThreePhase
 *    (source/QuadraticFreeEnergyStrategyMultiOrderTernaryThreePhase.cc
 *    :120-210, 237-520), added to the multi-order interface terms of
 *    MultiOrderRHSStrategy::evaluateRHS (source/MultiOrderRHSStrategy.cc
 *    :146-259) before the product by the mobility;
 *  - face diffusivities of three phases (TbasedCompositionDiffusionStrategy::
 *    setDiffusion, source/TbasedCompositionDiffusionStrategy.cc:225-300 ->
 *    concentration_pfmdiffusion_of_temperature_threephases, source/fortran/
 *    2d/concentrationdiffusion.m4:523-640; the ternary diagonal copies, :335)
 *    and the EBS composition RHS with two species (EBSCompositionRHSStrategy::
 *    computeFluxOnPatch, source/EBSCompositionRHSStrategy.cc:183-328 with
 *    add_flux, 2d/flux.m4:11-66, for L, A then B; computerhsconcentration).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R_GAS 8.314472 /* GASCONSTANT_R_JPKPMOL (Thermo4PFM; value unpinned) */

#define CI(i, j, k) \
   ((long)((i) + G->g) + (long)((j) + G->g) * G->sy + (long)((k) + G->gz) * G->sz)
#define FX(i, j, k) ((long)(i) + (long)(G->nx + 1) * ((long)(j) + (long)G->ny * (k)))
#define FY(i, j, k) ((long)(i) + (long)G->nx * ((long)(j) + (long)(G->ny + 1) * (k)))
#define FZ(i, j, k) ((long)(i) + (long)G->nx * ((long)(j) + (long)G->ny * (k)))
#define YOFF (G->nfx)
#define ZOFF (G->nfx + G->nfy)

/* phase weights of a cell (:142-157 of the KKS strategy, :470-493 of the
 * free energy strategy): [0] liquid = the last order, [1] solid A = orders
 * 0..norderp_A-1, [2] solid B = the rest; returns 1 / sum phi^2 */
static double tern3_weights(const ampe_model_params* p, int norder, const double* phi, long ng,
                            long c, double h[3])
{
   double hl = phi[(long)(norder - 1) * ng + c] * phi[(long)(norder - 1) * ng + c];
   double ha = 0.0, hb = 0.0;
   for (int i = 0; i < p->norderp_A; i++) ha += phi[(long)i * ng + c] * phi[(long)i * ng + c];
   for (int i = p->norderp_A; i < norder - 1; i++)
      hb += phi[(long)i * ng + c] * phi[(long)i * ng + c];
   double sum2 = hl + ha + hb;
   double inv = 1. / sum2;
   h[0] = hl * inv;
   h[1] = ha * inv;
   h[2] = hb * inv;
   return inv;
}

/* Thermo4PFM QuadraticFreeEnergyFunctionsTernaryThreePhase::
 * computePhaseConcentrations, restated in closed form: x = (cL0, cL1, cA0,
 * cA1, cB0, cB1) */
void o_tern3_phase_conc(const ampe_model_params* p, const double c[2], const double h[3],
                        double x[6])
{
   double g[3];
   for (int a = 0; a < 3; a++) g[a] = o_interp_func(h[a], p->conc_interp);
   for (int k = 0; k < 2; k++) {
      double num = c[k], den = 0.0;
      for (int a = 0; a < 3; a++) {
         num = num - g[a] * p->tern_ceq[a][k];
         den = den + g[a] / (2.0 * p->tern_A[a][k]);
      }
      double mu = num / den;
      for (int a = 0; a < 3; a++) x[2 * a + k] = p->tern_ceq[a][k] + mu / (2.0 * p->tern_A[a][k]);
   }
}

/* computePhaseConcentrationsOnPatch over the ghost box: ph holds six fields
 * (cL0, cL1, cA0, cA1, cB0, cB1), c two */
void o_kks_tern3(const OGeom* G, const ampe_model_params* p, int norder, const double* phi,
                 const double* c, double* ph)
{
   long ng = G->ncg;
   for (long n = 0; n < ng; n++) {
      double h[3], x[6];
      tern3_weights(p, norder, phi, ng, n, h);
      double cc[2] = {c[n], c[ng + n]};
      o_tern3_phase_conc(p, cc, h, x);
      for (int m = 0; m < 6; m++) ph[(long)m * ng + n] = x[m];
   }
}

/* free energy [pJ/um^3] and potentials of phase a at x (computeFreeEnergy,
 * computeMuL/A/B: the J/mol values times 1e-6 / v_m) */
static double tern3_energy(const ampe_model_params* p, int a, const double x[2], double mu[2])
{
   double vm = a == 0 ? p->vm_L : (a == 1 ? p->vm_A : p->vm_B);
   double fac = 1.e-6 / vm;
   double f = 0.0;
   for (int k = 0; k < 2; k++) {
      double d = x[k] - p->tern_ceq[a][k];
      f = f + p->tern_A[a][k] * d * d;
      mu[k] = 2.0 * p->tern_A[a][k] * d * fac;
   }
   return f * fac;
}

/* Moelans' driving force of order ip (addDrivingForceOnPatch, :462-510) */
static double tern3_driving(const ampe_model_params* p, int norder, const double* phi, long ng,
                            long c, const double* ph, int ip)
{
   double cl[2] = {ph[c], ph[ng + c]};
   double ca[2] = {ph[2 * ng + c], ph[3 * ng + c]};
   double cb[2] = {ph[4 * ng + c], ph[5 * ng + c]};
   double muL[2], muA[2], muB[2];
   double fl = tern3_energy(p, 0, cl, muL);
   double fa = tern3_energy(p, 1, ca, muA);
   double fb = tern3_energy(p, 2, cb, muB);
   double dfa = (fa - muA[0] * ca[0] - muA[1] * ca[1]);
   double dfb = (fb - muB[0] * cb[0] - muB[1] * cb[1]);
   double dfl = (fl - muL[0] * cl[0] - muL[1] * cl[1]);
   double h[3];
   double sum2inv = tern3_weights(p, norder, phi, ng, c, h);
   double hl = h[0], ha = h[1], hb = h[2];
   double v = phi[(long)ip * ng + c];
   if (ip < p->norderp_A) return 2. * v * ((1. - ha) * dfa - hl * dfl - hb * dfb) * sum2inv;
   if (ip < norder - 1) return 2. * v * ((1. - hb) * dfb - hl * dfl - ha * dfa) * sum2inv;
   return 2. * v * ((1. - hl) * dfl - ha * dfa - hb * dfb) * sum2inv;
}

double o_tern3_driving_force(const ampe_model_params* p, int norder, const double* phi_cell,
                             const double ph_cell[6], int ip)
{
   return tern3_driving(p, norder, phi_cell, 1, 0, ph_cell, ip);
}

/* MultiOrderRHSStrategy::evaluateRHS with the ternary three-phase free
 * energy: the interface terms of o_phase_rhs_multiorder, minus the driving
 * force, times the mobility (field M or the uniform phi_mobility) */
/* The three-phase phase RHS from given per-order side fluxes (norder blocks
 * of nf faces, the composite fluxes on a hierarchy). */
void o_phase_rhs_tern3_fx(const OGeom* G, const ampe_model_params* p, int norder,
                          const double* fl, const double* phi, const double* ph,
                          const double* M, double* rhs)
{
   long ng = G->ncg;
   ampe_model_params q = *p;
   q.phi_mobility = 1.0;
   o_phase_rhs_multiorder_fx(G, &q, norder, fl, phi, NULL, rhs);
   for (int ip = 0; ip < norder; ip++)
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               long c = CI(i, j, k);
               double r = rhs[(long)ip * ng + c];
               r = r - tern3_driving(p, norder, phi, ng, c, ph, ip);
               rhs[(long)ip * ng + c] = r * (M ? M[c] : p->phi_mobility);
            }
}

void o_phase_rhs_tern3(const OGeom* G, const ampe_model_params* p, int norder,
                       const double* phi, const double* ph, const double* M, double* rhs)
{
   long nf = G->nfx + G->nfy + G->nfz;
   double* fl = (double*)malloc(sizeof(double) * nf * norder);
   o_phase_flux_multiorder(G, p, norder, phi, fl);
   o_phase_rhs_tern3_fx(G, p, norder, fl, phi, ph, M, rhs);
   free(fl);
}

/* concentration_pfmdiffusion_of_temperature_threephases: D_a on every face,
 * a = L (the last order), A (sum of the face averages of its orders), B */
void o_ebs3_coeffs(const OGeom* G, const ampe_model_params* p, int norder, const double* phi,
                   double* DL, double* DA, double* DB)
{
   long st[3] = {1, G->sy, G->sz};
   long ng = G->ncg;
   double qlr = p->Q0_liquid / R_GAS, qar = p->Q0_solid / R_GAS, qbr = p->Q0_solid_B / R_GAS;
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - st[dir];
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               double vphi = 0.0;
               for (int ip = 0; ip < p->norderp_A; ip++)
                  vphi = vphi + o_average_func(phi[(long)ip * ng + cm], phi[(long)ip * ng + c0],
                                               p->avg_type);
               double vA = o_interp_func(vphi, p->diff_interp);
               vphi = 0.0;
               for (int ip = p->norderp_A; ip < norder - 1; ip++)
                  vphi = vphi + o_average_func(phi[(long)ip * ng + cm], phi[(long)ip * ng + c0],
                                               p->avg_type);
               double vB = o_interp_func(vphi, p->diff_interp);
               long l = (long)(norder - 1) * ng;
               vphi = o_average_func(phi[l + cm], phi[l + c0], p->avg_type);
               double vL = o_interp_func(vphi, p->diff_interp);
               double invT = 2.0 / (p->temperature + p->temperature);
               double dl = p->D_liquid * exp(-qlr * invT);
               double da = p->D_solid * exp(-qar * invT);
               double db = p->D_solid_B * exp(-qbr * invT);
               DL[fi] = vL * dl;
               DA[fi] = vA * da;
               DB[fi] = vB * db;
            }
   }
}

/* EBS composition RHS of two species: flux_k = (dxinv D_L dcL_k + dxinv D_A
 * dcA_k) + dxinv D_B dcB_k (add_flux for L, A then B; the zero off-diagonal
 * entries of the ternary diffusion matrices add nothing), then
 * rhs_k = M_c div flux_k (rhs holds two fields) */
/* The two species' side fluxes of the three-phase EBS form,
 * J_k = sum over L, A, B of D_a grad c_ak (2 blocks of nf faces). */
void o_ebs3_flux(const OGeom* G, const ampe_model_params* p, const double* ph, const double* DL,
                 const double* DA, const double* DB, double* fl2)
{
   (void)p;
   long st[3] = {1, G->sy, G->sz};
   long ng = G->ncg;
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   long nf = G->nfx + G->nfy + G->nfz;
   for (int sp = 0; sp < 2; sp++) {
      const double* cl = ph + (long)sp * ng;
      const double* ca = ph + (long)(2 + sp) * ng;
      const double* cb = ph + (long)(4 + sp) * ng;
      double* fl = fl2 + (long)sp * nf;
      for (int dir = 0; dir < G->ndim; dir++) {
         int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
         long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
         for (int k = 0; k < ez; k++)
            for (int j = 0; j < ey; j++)
               for (int i = 0; i < ex; i++) {
                  long c0 = CI(i, j, k), cm = c0 - st[dir];
                  long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
                  double f = 0.0;
                  f = f + dinv[dir] * DL[fi] * (cl[c0] - cl[cm]);
                  f = f + dinv[dir] * DA[fi] * (ca[c0] - ca[cm]);
                  f = f + dinv[dir] * DB[fi] * (cb[c0] - cb[cm]);
                  fl[fi] = f;
               }
      }
   }
}

/* M_c div J_k of the two species from given side fluxes. */
void o_conc_rhs_ebs3_fx(const OGeom* G, const ampe_model_params* p, const double* fl2,
                        double* rhs)
{
   long ng = G->ncg;
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   long nf = G->nfx + G->nfy + G->nfz;
   for (int sp = 0; sp < 2; sp++) {
      const double* fl = fl2 + (long)sp * nf;
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               double s = dinv[0] * (fl[FX(i + 1, j, k)] - fl[FX(i, j, k)]) +
                          dinv[1] * (fl[YOFF + FY(i, j + 1, k)] - fl[YOFF + FY(i, j, k)]);
               if (G->ndim == 3)
                  s = s + dinv[2] * (fl[ZOFF + FZ(i, j, k + 1)] - fl[ZOFF + FZ(i, j, k)]);
               rhs[(long)sp * ng + CI(i, j, k)] = p->conc_mobility * s;
            }
   }
}

void o_conc_rhs_ebs3(const OGeom* G, const ampe_model_params* p, const double* ph,
                     const double* DL, const double* DA, const double* DB, double* rhs)
{
   long nf = G->nfx + G->nfy + G->nfz;
   double* fl = (double*)malloc(sizeof(double) * nf * 2);
   o_ebs3_flux(G, p, ph, DL, DA, DB, fl);
   o_conc_rhs_ebs3_fx(G, p, fl, rhs);
   free(fl);
}
