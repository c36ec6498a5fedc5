/*
 * Oracle kernels: FP64 C restatement of the reference's per-patch Fortran and
 * C++ loops.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Floating-point expressions keep the reference's operation order (Fortran
 * evaluates left to right); build with -ffp-contract=off so no FMA is formed.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R_GAS 8.314472 /* GASCONSTANT_R_JPKPMOL (Thermo4PFM; value unpinned) */

/* Single precision literal 1.e-12 of anisotropic_gradient_flux promoted to
 * double (source/fortran/2d/quatrhs.m4:193, 3d/quatrhs.m4:184). */
#define ANISO_THRESHOLD ((double)1.e-12f)

void o_geom_init(OGeom* G, int ndim, int nx, int ny, int nz, int g,
                 const double dx[3])
{
   G->ndim = ndim;
   G->nx = nx;
   G->ny = ny;
   G->nz = (ndim == 3) ? nz : 1;
   G->g = g;
   G->gz = (ndim == 3) ? g : 0;
   G->sy = nx + 2 * g;
   G->sz = G->sy * (ny + 2 * g);
   G->ncg = G->sz * (G->nz + 2 * G->gz);
   G->nfx = (long)(nx + 1) * ny * G->nz;
   G->nfy = (long)nx * (ny + 1) * G->nz;
   G->nfz = (ndim == 3) ? (long)nx * ny * (G->nz + 1) : 0;
   G->dx[0] = dx[0];
   G->dx[1] = dx[1];
   G->dx[2] = (ndim == 3) ? dx[2] : 1.0;
   G->periodic = 1;
}

#define CI(i, j, k) \
   ((long)((i) + G->g) + (long)((j) + G->g) * G->sy + (long)((k) + G->gz) * G->sz)
#define FX(i, j, k) ((long)(i) + (long)(G->nx + 1) * ((long)(j) + (long)G->ny * (k)))
#define FY(i, j, k) ((long)(i) + (long)G->nx * ((long)(j) + (long)(G->ny + 1) * (k)))
#define FZ(i, j, k) ((long)(i) + (long)G->nx * ((long)(j) + (long)G->ny * (k)))
#define YOFF (G->nfx)
#define ZOFF (G->nfx + G->nfy)

/* ------------------------------------------------------------------------
 * functions.f helpers (source/fortran/functions.f:17-407).  Only the first
 * character of the type selects the branch, as in the reference.
 * ---------------------------------------------------------------------- */
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* interp_func, source/fortran/functions.f:17-92 */
double o_interp_func(double phi, char type)
{
   double phit;
   switch (type) {
      case 'q': phit = dmax(0.0, phi); return phit * phit;
      case 'p':
         phit = dmax(0.0, dmin(1.0, phi));
         return phit * phit * phit * (10.0 - 15.0 * phit + 6.0 * phit * phit);
      case 'h':
         phit = dmax(0.0, dmin(1.0, phi));
         return phit * phit * (3.0 - 2.0 * phit);
      case 'w': phit = dmax(0.0, phi); return phit * phit * (2.0 - phit);
      case 'l':
         phit = dmax(0.0, phi);
         phit = dmin(1.0, phit);
         return phit;
      case 'm':
         phit = dmax(0.0, dmin(1.0, phi));
         return phit * phit / (2.0 * phit * (phit - 1.0) + 1.0);
      case '3': phit = dmax(0.0, phi); return phit * phit * phit;
      case 's': return log(cosh(10.0 * phi)) / log(cosh(10.0));
      case 'c': return 1.0;
      default: return NAN;
   }
}

/* deriv_interp_func, source/fortran/functions.f:95-173 */
double o_deriv_interp_func(double phi, char type)
{
   double phit, tmp;
   switch (type) {
      case 'q': phit = dmax(0.0, phi); return 2.0 * phit;
      case 'p':
         phit = dmax(0.0, dmin(1.0, phi));
         return 30.0 * phit * phit * (1.0 - phit) * (1.0 - phit);
      case 'h':
         phit = dmax(0.0, dmin(1.0, phi));
         return 6.0 * phit * (1.0 - phit);
      case 'w': phit = dmax(0.0, phi); return phit * (4.0 - 3.0 * phit);
      case 'l': return (phi > 0.0 || phi < 1.0) ? 1.0 : 0.0;
      case 'm':
         phit = dmax(0.0, dmin(1.0, phi));
         tmp = 2.0 * phit * (phit - 1.0) + 1.0;
         return 2.0 * phit * (1.0 - phit) / (tmp * tmp);
      case '3': phit = dmax(0.0, phi); return 3.0 * phit * phit;
      case 's':
         phit = dmax(0.0, phi);
         return 10.0 * tanh(10.0 * phit) / log(cosh(10.0));
      case 'c': return 0.0;
      default: return NAN;
   }
}

/* second_deriv_interp_func, source/fortran/functions.f:176-243 */
double o_second_deriv_interp_func(double phi, char type)
{
   double phit;
   switch (type) {
      case 'q': return (phi >= 0.0) ? 2.0 : 0.0;
      case 'p':
         phit = dmax(0.0, dmin(1.0, phi));
         return 60.0 * phit * (1.0 - 3.0 * phit + 2.0 * phit * phit);
      case 'h':
         phit = dmax(0.0, dmin(1.0, phi));
         return 6.0 * (1.0 - 2.0 * phit);
      case 'w': phit = dmax(0.0, phi); return 4.0 - 6.0 * phit;
      case 'l': return 0.0;
      case 'c': return 0.0;
      default: return NAN;
   }
}

/* well_func & derivatives, source/fortran/functions.f:246-330 */
double o_well_func(double phi, char type)
{
   if (type == 'd') return 16.0 * phi * phi * (1.0 - phi) * (1.0 - phi);
   if (type == 's') return (1.0 - phi) * (1.0 - phi);
   return NAN;
}
double o_deriv_well_func(double phi, char type)
{
   if (type == 'd') return 32.0 * phi * (1.0 - phi) * (1.0 - 2.0 * phi);
   if (type == 's') return 2.0 * (phi - 1.0);
   return NAN;
}
double o_second_deriv_well_func(double phi, char type)
{
   if (type == 'd') return 32.0 * (1.0 + 6.0 * phi * (phi - 1.0));
   if (type == 's') return 2.0;
   return NAN;
}

/* average_func, source/fortran/functions.f:333-365 */
double o_average_func(double a, double b, char type)
{
   if (type == 'a') return 0.5 * (a + b);
   if (type == 'h') {
      if (a < 1.0e-16 || b < 1.0e-16) return 0.0;
      return 2.0 / (1.0 / a + 1.0 / b);
   }
   return NAN;
}

/* eval_grad_normi, source/fortran/quat.f:1497-1537 */
double o_eval_grad_normi(double gn2, char floor_type, double floor2,
                         double max_gni)
{
   if (floor_type == 'm') {
      if (gn2 > floor2) return pow(gn2, -0.5);
      return max_gni;
   }
   if (floor_type == 't') {
      double gng2 = gn2 * max_gni * max_gni;
      if (gng2 > 0.01) {
         double gn = sqrt(gn2);
         return tanh(max_gni * gn) / gn;
      }
      return max_gni * (1.0 - gng2 * (5.0 - 2.0 * gng2) / 15.0);
   }
   if (floor_type == 's') return pow(gn2 + floor2, -0.5);
   return NAN;
}

/* ------------------------------------------------------------------------
 * Free energies.
 * Quadratic (Thermo4PFM QuadraticFreeEnergyFunctionsBinary as configured by
 * source/QuadraticFreeEnergyStrategy.cc:30-77): f_i = A_i (c - ceq_i(T))^2,
 * ceq_i(T) = ceq_i + m_i (T - Tref), in J/mol.
 * CALPHAD binary (header of thermodynamic_data/calphadAuNi.dat:1-5,49-63):
 * G = c gA + (1-c) gB + RT[c ln c + (1-c) ln(1-c)] + c(1-c) sum L_i (2c-1)^i.
 * phase: 0 = liquid, 1 = solid A.
 * ---------------------------------------------------------------------- */
static double quad_ceq(const ampe_model_params* p, int ph)
{
   return p->quad_ceq[ph] + p->quad_m[ph] * (p->temperature - p->quad_Tref);
}

/* Dilute binary solution (restatement of the published dilute model; the
 * Thermo4PFM KKSFreeEnergyFunctionDiluteBinary energy is not available, parity
 * of the energy itself unpinned): f_L = RT (c ln c - c),
 * f_S = f_L - RT ln(k_e) c + dG, dG = RT (k_e - 1)(T - T_m)/m. */
static double dilute_dG(const ampe_model_params* p)
{
   double RT = R_GAS * p->temperature;
   return RT * (p->dilute_keq - 1.0) * (p->temperature - p->dilute_Tm) / p->dilute_m;
}

double o_fenergy(const ampe_model_params* p, double c, int ph)
{
   if (p->fe_model == AMPE_FE_QUADRATIC) {
      double d = c - quad_ceq(p, ph);
      return p->quad_A[ph] * d * d;
   }
   if (p->fe_model == AMPE_FE_KKS_DILUTE) {
      double RT = R_GAS * p->temperature;
      double f = RT * (c * log(c) - c);
      if (ph == 1) f = f - RT * log(p->dilute_keq) * c + dilute_dG(p);
      return f;
   }
   const ampe_calphad_phase* cp = &p->calphad[ph];
   double RT = R_GAS * p->temperature;
   double x = 2.0 * c - 1.0;
   double P = cp->L[0] + x * (cp->L[1] + x * (cp->L[2] + x * cp->L[3]));
   return c * cp->gA + (1.0 - c) * cp->gB +
          RT * (c * log(c) + (1.0 - c) * log(1.0 - c)) + c * (1.0 - c) * P;
}

double o_dfenergy(const ampe_model_params* p, double c, int ph)
{
   if (p->fe_model == AMPE_FE_QUADRATIC) {
      return 2.0 * p->quad_A[ph] * (c - quad_ceq(p, ph));
   }
   if (p->fe_model == AMPE_FE_KKS_DILUTE) {
      double RT = R_GAS * p->temperature;
      double mu = RT * log(c);
      if (ph == 1) mu = mu - RT * log(p->dilute_keq);
      return mu;
   }
   const ampe_calphad_phase* cp = &p->calphad[ph];
   double RT = R_GAS * p->temperature;
   double x = 2.0 * c - 1.0;
   double P = cp->L[0] + x * (cp->L[1] + x * (cp->L[2] + x * cp->L[3]));
   double dP = cp->L[1] + x * (2.0 * cp->L[2] + x * 3.0 * cp->L[3]);
   return cp->gA - cp->gB + RT * log(c / (1.0 - c)) +
          (1.0 - 2.0 * c) * P + 2.0 * c * (1.0 - c) * dP;
}

double o_d2fenergy(const ampe_model_params* p, double c, int ph)
{
   if (p->fe_model == AMPE_FE_QUADRATIC) return 2.0 * p->quad_A[ph];
   if (p->fe_model == AMPE_FE_KKS_DILUTE) return R_GAS * p->temperature / c;
   const ampe_calphad_phase* cp = &p->calphad[ph];
   double RT = R_GAS * p->temperature;
   double x = 2.0 * c - 1.0;
   double P = cp->L[0] + x * (cp->L[1] + x * (cp->L[2] + x * cp->L[3]));
   double dP = cp->L[1] + x * (2.0 * cp->L[2] + x * 3.0 * cp->L[3]);
   double d2P = 2.0 * cp->L[2] + x * 6.0 * cp->L[3];
   return RT / (c * (1.0 - c)) - 2.0 * P + 4.0 * (1.0 - 2.0 * c) * dP +
          4.0 * c * (1.0 - c) * d2P;
}

/* Phase mobility at one cell.  Constant M_phi (SimpleQuatMobilityStrategy)
 * or Kim's KKS mobility (KimMobilityStrategyInfMob::evaluateMobility,
 * source/KimMobilityStrategyInfMob.cc:54-83; FiniteMob, header :29-38;
 * FiniteMobAntiTrap :54-80) with the factor of kks_mobility_factor
 * (source/KKStools.cc:13-30): zeta = (cL - cS) f_L''(cL) (cL - cS) / D_L(T),
 * converted by 1e-6 / v_m; InfMob: factor / zeta; FiniteMob: 1 / alpha;
 * FiniteMobAntiTrap: 1 / (alpha + zeta / factor), alpha = 3 sqrt(2) xi / mu,
 * xi = eps / sqrt(16 w). */
double o_phase_mobility(const ampe_model_params* p, double cl, double cs)
{
   if (p->phi_mobility_type != AMPE_PHI_MOB_KIM) return p->phi_mobility;
   double eps = p->epsilon_phi, w = p->phi_well_scale;
   double a2 = (p->energy_interp == 'p') ? 47. / 60. : 0.5;
   double factor = 16. * w / (3. * eps * eps * a2);
   double alpha = 0.0;
   int finite = p->kim_interface_mobility > 0.0;
   if (finite) {
      double xi = eps / sqrt(16. * w);
      alpha = 3. * sqrt(2.) * xi / p->kim_interface_mobility;
      if (!p->kim_antitrapping) return 1. / alpha;
   }
   double T = p->temperature;
   double DL = p->D_liquid * exp(-p->Q0_liquid / (R_GAS * T));
   double d2 = o_d2fenergy(p, cl, 0);
   double zeta = 0.;
   zeta += (cl - cs) * d2 * (cl - cs);
   double zf = zeta / DL;
   zf *= (1.e-6 / p->vm_L);
   if (finite) return 1. / (alpha + (1. / factor) * zf);
   return factor / zf;
}

/* Per-cell KKS: find (cL, cS) with mu_L(cL) = mu_S(cS) and
 * (1-h) cL + h cS = c, h = interp(phi, conc_interp).
 * Quadratic wells: closed form (the linear 2x2 system).
 * CALPHAD: Newton on F = [(mu_L - mu_S)/RT, (1-h)cL + h cS - c] from the
 * guess in (*cl, *cs), step halving to stay inside (0,1).  Returns the
 * Newton iteration count, or -1 without convergence (reference:
 * source/CALPHADequilibriumPhaseConcentrationsStrategy.cc:343-402 calling
 * Thermo4PFM computePhaseConcentrations; iterate sequence unpinned). */
int o_kks_cell(const ampe_model_params* p, double c, double phi, double* cl,
               double* cs)
{
   double h = o_interp_func(phi, p->conc_interp);
   if (p->fe_model == AMPE_FE_QUADRATIC) {
      double eL = quad_ceq(p, 0), eS = quad_ceq(p, 1);
      double r = p->quad_A[0] / p->quad_A[1];
      double xl = (c - h * (eS - r * eL)) / (1.0 - h + h * r);
      *cl = xl;
      *cs = eS + r * (xl - eL);
      return 0;
   }
   if (p->fe_model == AMPE_FE_KKS_DILUTE) {
      /* mu_L = mu_S  <=>  cS = k_e cL; mass balance gives cL */
      double k = p->dilute_keq;
      double xl = c / (1.0 - h + h * k);
      *cl = xl;
      *cs = k * xl;
      return 0;
   }
   double RTinv = 1.0 / (R_GAS * p->temperature);
   double xl = *cl, xs = *cs;
   const double lo = 1.0e-12, hi = 1.0 - 1.0e-12;
   if (!(xl > lo)) xl = lo;
   if (!(xl < hi)) xl = hi;
   if (!(xs > lo)) xs = lo;
   if (!(xs < hi)) xs = hi;
   for (int it = 0; it <= p->newton_max_its; it++) {
      double f1 = (o_dfenergy(p, xl, 0) - o_dfenergy(p, xs, 1)) * RTinv;
      double f2 = (1.0 - h) * xl + h * xs - c;
      if (fabs(f1) < p->newton_tol && fabs(f2) < p->newton_tol) {
         *cl = xl;
         *cs = xs;
         return it;
      }
      if (it == p->newton_max_its) break;
      double a = o_d2fenergy(p, xl, 0) * RTinv;
      double b = -o_d2fenergy(p, xs, 1) * RTinv;
      double cc = 1.0 - h, d = h;
      double det = a * d - b * cc;
      double dl = -(d * f1 - b * f2) / det;
      double ds = -(a * f2 - cc * f1) / det;
      double alpha = 1.0;
      for (int k = 0; k < 60; k++) {
         double nl = xl + alpha * dl, ns = xs + alpha * ds;
         if (nl > 0.0 && nl < 1.0 && ns > 0.0 && ns < 1.0) break;
         alpha *= 0.5;
      }
      xl = xl + alpha * dl;
      xs = xs + alpha * ds;
   }
   *cl = xl;
   *cs = xs;
   return -1;
}

/* PhaseConcentrationsStrategy::computePhaseConcentrations over the ghost box
 * (source/PhaseConcentrationsStrategy.cc:26-127; ghost cells included per
 * source/QuatIntegrator.cc:3119-3126).  Guess from (clr, csr); returns the
 * largest iteration count or -1. */
int o_kks_solve(const OGeom* G, const ampe_model_params* p, const double* phi,
                const double* c, const double* clr, const double* csr,
                double* cl, double* cs)
{
   int maxit = 0, fail = 0;
#pragma omp parallel for reduction(max : maxit) reduction(| : fail)
   for (long n = 0; n < G->ncg; n++) {
      double xl = clr[n], xs = csr[n];
      int it = o_kks_cell(p, c[n], phi[n], &xl, &xs);
      cl[n] = xl;
      cs[n] = xs;
      if (it < 0) fail = 1;
      else if (it > maxit) maxit = it;
   }
   return fail ? -1 : maxit;
}

/* Phase mobility field over interior cells (QuatIntegrator::computeMobilities,
 * source/QuatIntegrator.cc:2959-2992, after the phase concentrations). */
void o_phase_mobility_field(const OGeom* G, const ampe_model_params* p, const double* cl,
                            const double* cs, double* M)
{
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            M[c0] = o_phase_mobility(p, cl[c0], cs[c0]);
         }
}

/* ------------------------------------------------------------------------
 * Periodic ghost fill of a single patch covering the whole domain (SAMRAI
 * periodic shift, closed form: ghost value = interior value of the wrapped
 * index).
 * ---------------------------------------------------------------------- */
static int wrap(int i, int n) { return ((i % n) + n) % n; }

void o_ghost_fill_periodic(const OGeom* G, double* f, int ncomp)
{
   int g = G->g, gz = G->gz;
   for (int m = 0; m < ncomp; m++) {
      double* a = f + (long)m * G->ncg;
      for (int k = -gz; k < G->nz + gz; k++)
         for (int j = -g; j < G->ny + g; j++)
            for (int i = -g; i < G->nx + g; i++) {
               if (i >= 0 && i < G->nx && j >= 0 && j < G->ny && k >= 0 &&
                   k < G->nz)
                  continue;
               a[CI(i, j, k)] =
                   a[CI(wrap(i, G->nx), wrap(j, G->ny), wrap(k, G->nz))];
            }
   }
}

static int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

void o_ghost_fill(const OGeom* G, double* f, int ncomp)
{
   if (G->periodic) {
      o_ghost_fill_periodic(G, f, ncomp);
      return;
   }
   int g = G->g, gz = G->gz;
   for (int m = 0; m < ncomp; m++) {
      double* a = f + (long)m * G->ncg;
      for (int k = -gz; k < G->nz + gz; k++)
         for (int j = -g; j < G->ny + g; j++)
            for (int i = -g; i < G->nx + g; i++) {
               if (i >= 0 && i < G->nx && j >= 0 && j < G->ny && k >= 0 &&
                   k < G->nz)
                  continue;
               a[CI(i, j, k)] =
                   a[CI(clampi(i, G->nx), clampi(j, G->ny), clampi(k, G->nz))];
            }
   }
}

/* ------------------------------------------------------------------------
 * Phase flux strategies.
 * ---------------------------------------------------------------------- */

/* gradient_flux: source/fortran/2d/quatrhs.m4:12-53, 3d/quatrhs.m4:12-62 */
static void gradient_flux(const OGeom* G, double eps, const double* ph,
                          double* flux)
{
   double e2 = eps * eps;
   double dxinv = e2 / G->dx[0], dyinv = e2 / G->dx[1], dzinv = e2 / G->dx[2];
#pragma omp parallel for
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i <= G->nx; i++)
            flux[FX(i, j, k)] = (ph[CI(i, j, k)] - ph[CI(i - 1, j, k)]) * dxinv;
#pragma omp parallel for
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j <= G->ny; j++)
         for (int i = 0; i < G->nx; i++)
            flux[YOFF + FY(i, j, k)] =
                (ph[CI(i, j, k)] - ph[CI(i, j - 1, k)]) * dyinv;
   if (G->ndim == 3) {
#pragma omp parallel for
      for (int k = 0; k <= G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++)
               flux[ZOFF + FZ(i, j, k)] =
                   (ph[CI(i, j, k)] - ph[CI(i, j, k - 1)]) * dzinv;
   }
}

/* compute_flux_isotropic (Shukla-Giri 9-point, 2-D only):
 * source/fortran/2d/quatrhs.m4:106-151 */
static void flux_isotropic_2d(const OGeom* G, double eps, const double* ph,
                              double* flux)
{
   double e2 = eps * eps;
   double dxinv = (1.0 / 12.0) * e2 / G->dx[0];
   double dyinv = (1.0 / 12.0) * e2 / G->dx[1];
   for (int j = 0; j < G->ny; j++)
      for (int i = 0; i <= G->nx; i++)
         flux[FX(i, j, 0)] =
             dxinv * ((ph[CI(i, j - 1, 0)] - ph[CI(i - 1, j - 1, 0)]) +
                      (ph[CI(i, j, 0)] - ph[CI(i - 1, j, 0)]) * 10.0 +
                      (ph[CI(i, j + 1, 0)] - ph[CI(i - 1, j + 1, 0)]));
   for (int j = 0; j <= G->ny; j++)
      for (int i = 0; i < G->nx; i++)
         flux[YOFF + FY(i, j, 0)] =
             dyinv * ((ph[CI(i - 1, j, 0)] - ph[CI(i - 1, j - 1, 0)]) +
                      (ph[CI(i, j, 0)] - ph[CI(i, j - 1, 0)]) * 10.0 +
                      (ph[CI(i + 1, j, 0)] - ph[CI(i + 1, j - 1, 0)]));
}

/* anisotropic_gradient_flux, 2-D: source/fortran/2d/quatrhs.m4:154-256 */
static void flux_aniso_2d(const OGeom* G, double eps, double nu, int kn,
                          const double* ph, const double* q1, int qlen,
                          double* flux)
{
   double pi = 4.0 * atan(1.0);
   double dxinv = 1.0 / G->dx[0], dyinv = 1.0 / G->dx[1];
   for (int j = 0; j < G->ny; j++)
      for (int i = 0; i <= G->nx; i++) {
         double dphidx = (ph[CI(i, j, 0)] - ph[CI(i - 1, j, 0)]) * dxinv;
         double dphidy = 0.25 * (ph[CI(i - 1, j + 1, 0)] - ph[CI(i - 1, j - 1, 0)] +
                                 ph[CI(i, j + 1, 0)] - ph[CI(i, j - 1, 0)]) *
                         dyinv;
         double theta =
             (fabs(dphidx) > ANISO_THRESHOLD) ? atan(dphidy / dphidx) : 0.5 * pi;
         double q = 0.5 * (q1[CI(i - 1, j, 0)] + q1[CI(i, j, 0)]);
         if (q > 1.0) q = 1.0;
         if (q < -1.0) q = -1.0;
         double phi = (qlen == 4) ? 2.0 * acos(q) : acos(q);
         double epst = eps * (1.0 + nu * cos(kn * (theta - phi)));
         double deps = -kn * eps * nu * sin(kn * (theta - phi));
         flux[FX(i, j, 0)] = epst * epst * dphidx - epst * deps * dphidy;
      }
   for (int j = 0; j <= G->ny; j++)
      for (int i = 0; i < G->nx; i++) {
         double dphidx = 0.25 * (ph[CI(i + 1, j - 1, 0)] - ph[CI(i - 1, j - 1, 0)] +
                                 ph[CI(i + 1, j, 0)] - ph[CI(i - 1, j, 0)]) *
                         dxinv;
         double dphidy = (ph[CI(i, j, 0)] - ph[CI(i, j - 1, 0)]) * dyinv;
         double theta =
             (fabs(dphidx) > ANISO_THRESHOLD) ? atan(dphidy / dphidx) : 0.5 * pi;
         double q = 0.5 * (q1[CI(i, j - 1, 0)] + q1[CI(i, j, 0)]);
         if (q > 1.0) q = 1.0;
         if (q < -1.0) q = -1.0;
         double phi = (qlen == 4) ? 2.0 * acos(q) : acos(q);
         double epst = eps * (1.0 + nu * cos(kn * (theta - phi)));
         double deps = -kn * eps * nu * sin(kn * (theta - phi));
         flux[YOFF + FY(i, j, 0)] = epst * epst * dphidy + epst * deps * dphidx;
      }
}

/* quatmult4 / quatconj: source/fortran/quat.f:867-927 */
static void quatmult4(const double* a, const double* b, double* r)
{
   r[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
   r[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
   r[2] = a[0] * b[2] + a[2] * b[0] + a[3] * b[1] - a[1] * b[3];
   r[3] = a[0] * b[3] + a[3] * b[0] + a[1] * b[2] - a[2] * b[1];
}

/* compute_dgamma: source/fortran/3d/quatrhs.m4:69-95 */
static void compute_dgamma(const double* q, const double* n, double* dg,
                           double* n4)
{
   double qp[4] = {q[0], -q[1], -q[2], -q[3]}, t[4], np[4];
   quatmult4(n, qp, t);
   quatmult4(q, t, np);
   /* x**4 evaluated as (x*x)*(x*x), as Fortran compilers expand it */
   double a1 = np[1] * np[1], a2 = np[2] * np[2], a3 = np[3] * np[3];
   *n4 = a1 * a1 + a2 * a2 + a3 * a3;
   double d[4];
   d[0] = 0.0;
   d[1] = np[1] * (np[1] * np[1] - *n4);
   d[2] = np[2] * (np[2] * np[2] - *n4);
   d[3] = np[3] * (np[3] * np[3] - *n4);
   quatmult4(d, q, t);
   quatmult4(qp, t, dg);
}

/* anisotropic_gradient_flux, 3-D (cubic): source/fortran/3d/quatrhs.m4:149-356.
 * dir: face direction; (a,b) the transverse directions. */
static void flux_aniso_3d(const OGeom* G, double eps, double eps4,
                          const double* ph, const double* q, double* flux)
{
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   double factor = 4.0 * eps4 / (1.0 - 3.0 * eps4);
   long cs[3] = {1, G->sy, G->sz};
   for (int dir = 0; dir < 3; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      double* fl = flux + (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
#pragma omp parallel for
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - cs[dir];
               double gr[3];
               for (int d = 0; d < 3; d++) {
                  if (d == dir)
                     gr[d] = (ph[c0] - ph[cm]) * dinv[d];
                  else
                     gr[d] = 0.25 *
                             (ph[cm + cs[d]] - ph[cm - cs[d]] + ph[c0 + cs[d]] -
                              ph[c0 - cs[d]]) *
                             dinv[d];
               }
               double gphi2 = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
               double dg[4], n4;
               if (fabs(gphi2) > ANISO_THRESHOLD) {
                  double nni = 1.0 / sqrt(gphi2);
                  double n[4] = {0.0, gr[0] * nni, gr[1] * nni, gr[2] * nni};
                  double qa[4];
                  for (int m = 0; m < 4; m++)
                     qa[m] = 0.5 * (q[(long)m * G->ncg + cm] + q[(long)m * G->ncg + c0]);
                  compute_dgamma(qa, n, dg, &n4);
               } else {
                  dg[0] = 0.0;
                  dg[1] = 0.0;
                  dg[2] = (dir == 2) ? 1.0 : 0.0;
                  dg[3] = (dir == 2) ? 0.0 : 1.0;
                  n4 = 0.0;
               }
               double gamma = eps * (1.0 - 3.0 * eps4) * (1.0 + factor * n4);
               long fi = (long)i + (long)ex * ((long)j + (long)ey * k);
               fl[fi] = gamma * gamma * gr[dir] +
                        16.0 * eps * gamma * eps4 * sqrt(gphi2) * dg[dir + 1];
            }
   }
}

/* PhaseFluxStrategy factory dispatch (source/PhaseFluxStrategyFactory.h:16-33) */
void o_phase_flux(const OGeom* G, const ampe_model_params* p,
                  const double* phi, const double* q, double* flux)
{
   if (p->flux_type == AMPE_FLUX_ANISOTROPIC) {
      if (G->ndim == 2)
         flux_aniso_2d(G, p->epsilon_phi, p->aniso_nu, p->aniso_knumber, phi, q,
                       p->qlen, flux);
      else
         flux_aniso_3d(G, p->epsilon_phi, p->aniso_nu, phi, q, flux);
   } else if (p->flux_type == AMPE_FLUX_ISOTROPIC && G->ndim == 2) {
      flux_isotropic_2d(G, p->epsilon_phi, phi, flux);
   } else {
      gradient_flux(G, p->epsilon_phi, phi, flux);
   }
}

/* ------------------------------------------------------------------------
 * Quaternion gradients.
 * ---------------------------------------------------------------------- */

/* |grad q| from cell-centred gradients (default quat_grad_modulus_type
 * "cells", source/QuatModelParameters.cc:692; QuatIntegrator.cc:2812-2817):
 * quatdiffs (2d/quatdiffs.m4:56-100) -> quatgrad_cell (2d/quatgrad.m4:12-60,
 * 3d/quatgrad.m4:12-70) -> quatgrad_modulus (2d/quatgrad.m4:882-915,
 * 3d quatgrad_modulus).  Interior cells. */
void o_quat_grad_modulus(const OGeom* G, const ampe_model_params* p,
                         const double* q, double* gmod)
{
   double p5[3] = {0.5 / G->dx[0], 0.5 / G->dx[1], 0.5 / G->dx[2]};
   long cs[3] = {1, G->sy, G->sz};
#pragma omp parallel for
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            double s = 0.0;
            for (int m = 0; m < p->qlen; m++) {
               const double* a = q + (long)m * G->ncg;
               double gx = ((a[c0 + cs[0]] - a[c0]) + (a[c0] - a[c0 - cs[0]])) * p5[0];
               double gy = ((a[c0 + cs[1]] - a[c0]) + (a[c0] - a[c0 - cs[1]])) * p5[1];
               if (G->ndim == 3) {
                  double gz = ((a[c0 + cs[2]] - a[c0]) + (a[c0] - a[c0 - cs[2]])) * p5[2];
                  s = s + gx * gx + gy * gy + gz * gz;
               } else {
                  s = s + gx * gx + gy * gy;
               }
            }
            gmod[c0] = sqrt(s);
         }
}

/* |grad q|^2 on every face from the full side gradient (normal component
 * dq/dh, transverse components as 4-point averages of differences):
 * quatgrad_side (2d/quatgrad.m4:157-229, 3d/quatgrad.m4:213) summed as in
 * compute_face_coef2d/3d (2d/quatfacops.m4:83-92): direction n outer,
 * component m inner. */
void o_quat_face_norm2(const OGeom* G, const ampe_model_params* p,
                       const double* q, double* n2)
{
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   double p25[3] = {0.25 * dinv[0], 0.25 * dinv[1], 0.25 * dinv[2]};
   long cs[3] = {1, G->sy, G->sz};
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      double* out = n2 + (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
#pragma omp parallel for
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - cs[dir];
               double s = 0.0;
               for (int n = 0; n < G->ndim; n++)
                  for (int m = 0; m < p->qlen; m++) {
                     const double* a = q + (long)m * G->ncg;
                     double gv;
                     if (n == dir) {
                        gv = dinv[n] * (a[c0] - a[cm]);
                     } else {
                        long t = cs[n];
                        /* diff_n(x) = a(x) - a(x - e_n); the 4 differences
                         * straddling the face, in the reference's order */
                        gv = p25[n] * ((a[cm + t] - a[cm]) + (a[cm] - a[cm - t]) +
                                       (a[c0 + t] - a[c0]) + (a[c0] - a[c0 - t]));
                     }
                     s = s + gv * gv;
                  }
               out[(long)i + (long)ex * ((long)j + (long)ey * k)] = s;
            }
   }
}

/* quatmobility with ghost cells: source/fortran/2d/mobility.m4:13-90 */
void o_quat_mobility(const OGeom* G, const ampe_model_params* p,
                     const double* phi, double* mq)
{
#pragma omp parallel for
   for (long n = 0; n < G->ncg; n++) {
      double ph = phi[n], qf;
      if (p->quat_mobility_type == 'e') {
         double c = p->quat_mobility_scale;
         ph = dmax(0.0, dmin(1.0, ph));
         qf = (1.0 - exp(c * ph)) / (1.0 - exp(c));
         qf = 1.0 - qf;
      } else if (p->quat_mobility_type == 'i') {
         ph = dmax(1.0e-6, dmin(1.0, ph));
         qf = dmax(0.0, (1.0 - ph) / (ph * ph));
         qf = dmin(qf, p->quat_mobility_scale);
      } else {
         ph = dmax(0.0, dmin(1.0, ph));
         qf = ph * ph * ph * (10.0 - 15.0 * ph + 6.0 * ph * ph);
         qf = 1.0 - qf;
      }
      mq[n] = p->min_quat_mobility + (p->quat_mobility - p->min_quat_mobility) * qf;
   }
}

/* ------------------------------------------------------------------------
 * Phase RHS: PhaseRHSStrategyWithQ::evaluateRHS per patch
 * (source/PhaseRHSStrategyWithQ.cc:151-312): flux, COMPUTERHSPBG
 * (2d/quatrhs.m4:264-405, 3d/quatrhs.m4:357), addDrivingForce
 * (QuadraticFreeEnergyStrategy.cc:393-511 / CALPHADFreeEnergyStrategyBinary.cc
 * :521-651), multiply by the uniform phase mobility (:297).
 * ---------------------------------------------------------------------- */
void o_phase_rhs(const OGeom* G, const ampe_model_params* p,
                 const double* phi, const double* q, const double* cl,
                 const double* cs, double* rhs)
{
   double* flux = (double*)malloc(sizeof(double) * (G->nfx + G->nfy + G->nfz));
   o_phase_flux(G, p, phi, q, flux);
   o_phase_rhs_fx(G, p, flux, phi, q, cl, cs, rhs);
   free(flux);
}

/* computerhspbg from a given side flux (the composite flux of a level whose
 * coarse-fine faces hold the coarsened fine fluxes,
 * source/PhaseRHSStrategyWithQ.cc:126-144). */
void o_phase_rhs_fx(const OGeom* G, const ampe_model_params* p, const double* flux,
                    const double* phi, const double* q, const double* cl,
                    const double* cs, double* rhs)
{
   double* gmod = NULL;
   if (p->evolve_quat) {
      gmod = (double*)calloc(G->ncg, sizeof(double));
      o_quat_grad_modulus(G, p, q, gmod);
   }
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   double mf = 2.0 * p->H_parameter, T = p->temperature;
   double eq2 = 0.5 * p->epsilon_q * p->epsilon_q;
   double convL = 1.0e-6 / p->vm_L, convA = 1.0e-6 / p->vm_A;
#pragma omp parallel for
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            double dtx = (flux[FX(i + 1, j, k)] - flux[FX(i, j, k)]) * dinv[0];
            double dty = (flux[YOFF + FY(i, j + 1, k)] - flux[YOFF + FY(i, j, k)]) * dinv[1];
            double r = dtx + dty;
            if (G->ndim == 3) {
               double dtz = (flux[ZOFF + FZ(i, j, k + 1)] - flux[ZOFF + FZ(i, j, k)]) * dinv[2];
               r = r + dtz;
            }
            double ph = phi[c0];
            r = r - p->phi_well_scale * o_deriv_well_func(ph, p->phi_well_type);
            if (p->evolve_quat) {
               double p1 = o_deriv_interp_func(ph, p->orient_interp1);
               double p2 = o_deriv_interp_func(ph, p->orient_interp2);
               double gm = gmod[c0];
               r = r - mf * T * p1 * gm - p2 * eq2 * gm * gm;
            }
            if (p->with_conc) {
               double c_l = cl[c0], c_a = cs[c0];
               double f_l = o_fenergy(p, c_l, 0) * convL;
               double f_a = o_fenergy(p, c_a, 1) * convA;
               double mu = (p->fe_model == AMPE_FE_QUADRATIC)
                               ? o_dfenergy(p, c_l, 0) * convL
                               : o_dfenergy(p, c_a, 1) * convA;
               double hp = o_deriv_interp_func(ph, p->energy_interp);
               r += hp * ((f_l - f_a) - mu * (c_l - c_a));
            }
            double mob = (p->phi_mobility_type == AMPE_PHI_MOB_KIM)
                             ? o_phase_mobility(p, cl[c0], cs[c0])
                             : p->phi_mobility;
            rhs[c0] = mob * r;
         }
   free(gmod);
}

/* ------------------------------------------------------------------------
 * Quaternion RHS.
 * ---------------------------------------------------------------------- */

/* compute_face_coef2d/3d (2d/quatfacops.m4:14-122) with the face |grad q|^2
 * precomputed: fc = -grad_normi * 2H T p1(phi_avg) - eps_q^2 p2(phi_avg) */
void o_quat_face_coef(const OGeom* G, const ampe_model_params* p,
                      const double* phi, const double* n2, double* fc)
{
   long cs[3] = {1, G->sy, G->sz};
   double floor2 = p->gradient_floor * p->gradient_floor;
   double maxgni = 1.0 / p->gradient_floor;
   double eps2 = p->epsilon_q * p->epsilon_q;
   double mf = 2.0 * p->H_parameter;
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
#pragma omp parallel for
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - cs[dir];
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               double phia = o_average_func(phi[cm], phi[c0], p->avg_type);
               double tempa = 0.5 * (p->temperature + p->temperature);
               double diff = mf * tempa * o_interp_func(phia, p->orient_interp1);
               double hphi2 = o_interp_func(phia, p->orient_interp2);
               double gni = o_eval_grad_normi(n2[fi], p->floor_type, floor2, maxgni);
               fc[fi] = -gni * diff - eps2 * hphi2;
            }
   }
}

/* QuatFACOps::evaluateRHS (source/QuatFACOps.cc:1861-1890) with
 * use_gradq_for_flux (QuatIntegrator.cc:2704-2708) and projection
 * (accumulateOperatorOnLevel, QuatFACOps.cc:2089-2108):
 * flux = fc * dq/dh (compute_flux2d_from_gradq, 2d/quatfacops.m4:232-282),
 * lambda (compute_lambda_flux2d :683-735, 3d analogue),
 * rhs = -M_q (div + 2 q lambda) (add_quat_proj_op2d :542-600).
 * n2: face |grad q|^2 used for the face coefficient (lagged copy). */
void o_quat_rhs(const OGeom* G, const ampe_model_params* p, const double* phi,
                const double* q, const double* n2, double* rhs)
{
   long nf = G->nfx + G->nfy + G->nfz;
   double* fl = (double*)malloc(sizeof(double) * nf * p->qlen);
   o_quat_flux(G, p, phi, q, n2, fl);
   o_quat_rhs_fx(G, p, phi, q, fl, rhs);
   free(fl);
}

/* compute_flux2d/3d: face flux fc * dq_m/dx of every component, fl[m][face] */
void o_quat_flux(const OGeom* G, const ampe_model_params* p, const double* phi,
                 const double* q, const double* n2, double* fl)
{
   long nf = G->nfx + G->nfy + G->nfz;
   double* fc = (double*)malloc(sizeof(double) * nf);
   o_quat_face_coef(G, p, phi, n2, fc);
   long cs[3] = {1, G->sy, G->sz};
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   for (int m = 0; m < p->qlen; m++) {
      const double* a = q + (long)m * G->ncg;
      for (int dir = 0; dir < G->ndim; dir++) {
         int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
         long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
#pragma omp parallel for
         for (int k = 0; k < ez; k++)
            for (int j = 0; j < ey; j++)
               for (int i = 0; i < ex; i++) {
                  long c0 = CI(i, j, k), cm = c0 - cs[dir];
                  long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
                  double gnormal = dinv[dir] * (a[c0] - a[cm]);
                  fl[(long)m * nf + fi] = fc[fi] * gnormal;
               }
      }
   }
   free(fc);
}

/* add_quat_op + add_quat_proj_op from given component fluxes fl[m][face] */
void o_quat_rhs_fx(const OGeom* G, const ampe_model_params* p, const double* phi,
                   const double* q, const double* fl, double* rhs)
{
   long nf = G->nfx + G->nfy + G->nfz;
   double* mq = (double*)malloc(sizeof(double) * G->ncg);
   o_quat_mobility(G, p, phi, mq);
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   double fac[3] = {0.5 / G->dx[0], 0.5 / G->dx[1], 0.5 / G->dx[2]};
#pragma omp parallel for
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            double lam = 0.0, sumq2 = 0.0;
            double dv[4];
            for (int m = 0; m < p->qlen; m++) {
               const double* f = fl + (long)m * nf;
               double left = f[FX(i, j, k)], right = f[FX(i + 1, j, k)];
               double down = f[YOFF + FY(i, j, k)], up = f[YOFF + FY(i, j + 1, k)];
               double qm = q[(long)m * G->ncg + c0];
               if (G->ndim == 3) {
                  double front = f[ZOFF + FZ(i, j, k)], back = f[ZOFF + FZ(i, j, k + 1)];
                  lam = lam - qm * ((right - left) * fac[0] + (up - down) * fac[1] +
                                    (back - front) * fac[2]);
                  dv[m] = (right - left) * dinv[0] + (up - down) * dinv[1] +
                          (back - front) * dinv[2];
               } else {
                  lam = lam - qm * ((right - left) * fac[0] + (up - down) * fac[1]);
                  dv[m] = (right - left) * dinv[0] + (up - down) * dinv[1];
               }
               sumq2 = sumq2 + qm * qm;
            }
            lam = lam / sumq2;
            for (int m = 0; m < p->qlen; m++) {
               double qm = q[(long)m * G->ncg + c0];
               rhs[(long)m * G->ncg + c0] = 0.0 - mq[c0] * (dv[m] + 2.0 * qm * lam);
            }
         }
   free(mq);
}

/* project2d/3d: source/fortran/2d/quatfacops.m4:940-997 (interior cells) */
void o_quat_project(const OGeom* G, const ampe_model_params* p,
                    const double* q, double* corr, double* err)
{
   int ql = p->qlen;
#pragma omp parallel for
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            double fac = 0.0;
            for (int m = 0; m < ql; m++) {
               double v = q[(long)m * G->ncg + c0];
               fac = fac + v * v;
            }
            fac = 1.0 / sqrt(fac);
            for (int m = 0; m < ql; m++)
               corr[(long)m * G->ncg + c0] = q[(long)m * G->ncg + c0] * fac;
            fac = 0.0;
            for (int m = 0; m < ql; m++)
               fac = fac + corr[(long)m * G->ncg + c0] * err[(long)m * G->ncg + c0];
            for (int m = 0; m < ql; m++)
               err[(long)m * G->ncg + c0] =
                   err[(long)m * G->ncg + c0] - corr[(long)m * G->ncg + c0] * fac;
            for (int m = 0; m < ql; m++)
               corr[(long)m * G->ncg + c0] =
                   corr[(long)m * G->ncg + c0] - q[(long)m * G->ncg + c0];
         }
}

/* ------------------------------------------------------------------------
 * Composition.
 * ---------------------------------------------------------------------- */

/* concentration_pfmdiffusion (2d/concentrationdiffusion.m4:14-96) and
 * KKSCompositionRHSStrategy::setDiffCoeffForPhaseOnPatch
 * (source/KKSCompositionRHSStrategy.cc:210-373). */
void o_conc_coeffs(const OGeom* G, const ampe_model_params* p,
                   const double* phi, const double* cl, const double* cs,
                   double* Dc, double* Dphi)
{
   long st[3] = {1, G->sy, G->sz};
   double qlr = p->Q0_liquid / R_GAS, qsr = p->Q0_solid / R_GAS;
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
#pragma omp parallel for
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - st[dir];
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               double vphi = o_average_func(phi[cm], phi[c0], p->avg_type);
               double hphi = o_interp_func(vphi, p->energy_interp);
               double invT = 2.0 / (p->temperature + p->temperature);
               double dl = p->D_liquid * exp(-qlr * invT);
               double ds = p->D_solid * exp(-qsr * invT);
               double d = (1.0 - hphi) * dl + hphi * ds;
               Dc[fi] = d;
               double ph = o_average_func(phi[c0], phi[cm], p->avg_type);
               double c_l = 0.5 * (cl[c0] + cl[cm]);
               double c_a = 0.5 * (cs[c0] + cs[cm]);
               double hp = o_deriv_interp_func(ph, p->energy_interp);
               Dphi[fi] = d * hp * (c_l - c_a);
            }
   }
}

/* concentrationflux (2d/concentrationrhs.m4:15-82, 3d :14) and
 * computerhsconcentration (2d/concentrationrhs.m4:326-367, 3d :412). */
void o_conc_rhs(const OGeom* G, const ampe_model_params* p, const double* c,
                const double* phi, const double* Dc, const double* Dphi,
                double* rhs)
{
   long nf = G->nfx + G->nfy + G->nfz;
   double* fl = (double*)malloc(sizeof(double) * nf);
   o_conc_flux(G, p, c, phi, Dc, Dphi, fl);
   o_conc_rhs_fx(G, p, fl, rhs);
   free(fl);
}

/* concentrationflux: J = (D_c dc + D_phi dphi) / dx on every face */
void o_conc_flux(const OGeom* G, const ampe_model_params* p, const double* c,
                 const double* phi, const double* Dc, const double* Dphi, double* fl)
{
   (void)p;
   long st[3] = {1, G->sy, G->sz};
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
#pragma omp parallel for
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - st[dir];
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               fl[fi] = dinv[dir] * (Dc[fi] * (c[c0] - c[cm]) +
                                     Dphi[fi] * (phi[c0] - phi[cm]));
            }
   }
}

/* computerhsconcentration from a given flux */
void o_conc_rhs_fx(const OGeom* G, const ampe_model_params* p, const double* fl, double* rhs)
{
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
#pragma omp parallel for
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            double s = dinv[0] * (fl[FX(i + 1, j, k)] - fl[FX(i, j, k)]) +
                       dinv[1] * (fl[YOFF + FY(i, j + 1, k)] - fl[YOFF + FY(i, j, k)]);
            if (G->ndim == 3)
               s = s + dinv[2] * (fl[ZOFF + FZ(i, j, k + 1)] - fl[ZOFF + FZ(i, j, k)]);
            rhs[CI(i, j, k)] = p->conc_mobility * s;
         }
}

/* EBS face diffusivities, two phases, uniform temperature:
 * concentration_pfmdiffusion_of_temperature (2d/concentrationdiffusion.m4
 * :341-428) as called by TbasedCompositionDiffusionStrategy::setDiffusion
 * (source/TbasedCompositionDiffusionStrategy.cc:305-322). */
void o_ebs_coeffs(const OGeom* G, const ampe_model_params* p, const double* phi, double* DL,
                  double* DA)
{
   long st[3] = {1, G->sy, G->sz};
   double qlr = p->Q0_liquid / R_GAS, qsr = p->Q0_solid / R_GAS;
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - st[dir];
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               double vphi = o_average_func(phi[cm], phi[c0], p->avg_type);
               double hphi = o_interp_func(vphi, p->diff_interp);
               double invT = 2.0 / (p->temperature + p->temperature);
               double dl = p->D_liquid * exp(-qlr * invT);
               double ds = p->D_solid * exp(-qsr * invT);
               DL[fi] = (1.0 - hphi) * dl;
               DA[fi] = hphi * ds;
            }
   }
}

/* CALPHADMobility::getQ (source/CALPHADMobility.h:42-55): value + R T ln(prefactor),
 * plus the T ln T and 1/T terms of a four-value entry. */
static double o_mob_q(const double q[2], const double qx[2], int has_extra, double T)
{
   double val = q[0] + R_GAS * T * log(q[1]);
   if (has_extra) val += qx[0] * T * log(T) + qx[1] / T;
   return val;
}

/* CALPHADMobility::getAtomicMobility with the binary getDeltaG
 * (source/CALPHADMobility.h:105-150, .cc:155-166): exp(dG / RT) / RT. */
double o_atomic_mobility(const ampe_calphad_mobility* m, double c0, double c1, double T)
{
   const double qa = o_mob_q(m->q[0], m->qx[0], m->has_extra, T);
   const double qb = o_mob_q(m->q[1], m->qx[1], m->has_extra, T);
   double qq[4];
   for (int n = 0; n < 4; n++) qq[n] = m->qq[n][0] + R_GAS * T * log(m->qq[n][1]);
   const double dc = c0 - c1;
   const double dG = c0 * qa + c1 * qb + c0 * c1 * (qq[0] + dc * (qq[1] + dc * (qq[2] + dc * qq[3])));
   const double rtinv = 1.0 / (R_GAS * T);
   return exp(dG * rtinv) * rtinv;
}

/* computeDiffusionMobilityBinaryPhase (source/CALPHADMobility.cc:200-218):
 * c0 c1 (c0 m1 + c1 m0) in um^2, m_k the atomic mobility of species k. */
double o_phase_diffusion_mobility(const ampe_calphad_mobility m[2], double c0, double T)
{
   const double c1 = 1.0 - c0;
   const double m0 = o_atomic_mobility(&m[0], c0, c1, T);
   const double m1 = o_atomic_mobility(&m[1], c0, c1, T);
   const double mm = c0 * m1 + c1 * m0;
   return c0 * c1 * mm * 1.0e12;
}

/* MobilityCompositionDiffusionStrategy::setDiffusion
 * (source/MobilityCompositionDiffusionStrategy.cc:97-150): per face the phase
 * concentrations and the temperature averaged arithmetically
 * (setDiffCoeffInEachPhaseOnPatch, :194-329), each phase's mobility times
 * d2f/dc2 in J/mol (computeLocalDiffusionMatrixL/A, :57-95), then the phase
 * weights (setPFMDiffOnPatch, :331-611). */
void o_ebs_coeffs_mob(const OGeom* G, const ampe_model_params* p, const double* phi,
                      const double* cl, const double* cs, double* DL, double* DA)
{
   long st[3] = {1, G->sy, G->sz};
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - st[dir];
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               double T = 0.5 * (p->temperature + p->temperature);
               double cL = 0.5 * (cl[c0] + cl[cm]), cA = 0.5 * (cs[c0] + cs[cm]);
               double dl = o_phase_diffusion_mobility(p->mobility[0], cL, T) *
                           o_d2fenergy(p, cL, 0);
               double da = o_phase_diffusion_mobility(p->mobility[1], cA, T) *
                           o_d2fenergy(p, cA, 1);
               double vphi = o_average_func(phi[cm], phi[c0], p->avg_type);
               double hphi = o_interp_func(vphi, p->diff_interp);
               DL[fi] = (1.0 - hphi) * dl;
               DA[fi] = hphi * da;
            }
   }
}

/* EBSCompositionRHSStrategy::computeFluxOnPatch (source/EBSCompositionRHSStrategy.cc
 * :183-328) with add_flux (2d/flux.m4:11-66, the liquid then the solid
 * contribution added to a zeroed flux) and computerhsconcentration
 * (2d/concentrationrhs.m4:326-367). */
void o_conc_rhs_ebs(const OGeom* G, const ampe_model_params* p, const double* cl,
                    const double* cs, const double* DL, const double* DA, double* rhs)
{
   long st[3] = {1, G->sy, G->sz};
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   long nf = G->nfx + G->nfy + G->nfz;
   double* fl = (double*)malloc(sizeof(double) * nf);
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
               f = f + dinv[dir] * DA[fi] * (cs[c0] - cs[cm]);
               fl[fi] = f;
            }
   }
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            double s = dinv[0] * (fl[FX(i + 1, j, k)] - fl[FX(i, j, k)]) +
                       dinv[1] * (fl[YOFF + FY(i, j + 1, k)] - fl[YOFF + FY(i, j, k)]);
            if (G->ndim == 3)
               s = s + dinv[2] * (fl[ZOFF + FZ(i, j, k + 1)] - fl[ZOFF + FZ(i, j, k)]);
            rhs[CI(i, j, k)] = p->conc_mobility * s;
         }
}

/* The same with add_flux_4th (2d/flux.m4:139-242, 3d/flux.m4:235-420):
 * interior faces add d (15/12 (c_h - c_m) - 1/12 (c_hh - c_mm)) / dx per
 * phase, the first and last face of a direction on a physical boundary
 * (physb, source/EBSCompositionRHSStrategy.cc:236-255; every side of the
 * single non-periodic patch) add d (c_h - c_m) / dx.  Needs G->g >= 2. */
void o_conc_rhs_ebs_4th(const OGeom* G, const ampe_model_params* p, const double* cl,
                        const double* cs, const double* DL, const double* DA, double* rhs)
{
   long st[3] = {1, G->sy, G->sz};
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   long nf = G->nfx + G->nfy + G->nfz;
   double* fl = (double*)malloc(sizeof(double) * nf);
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      int nd = dir == 0 ? G->nx : dir == 1 ? G->ny : G->nz;
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
      double d15 = 15.0 / (12.0 * G->dx[dir]), d1 = 1.0 / (12.0 * G->dx[dir]);
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               int id = dir == 0 ? i : dir == 1 ? j : k;
               int bnd = !G->periodic && (id == 0 || id == nd);
               long c0 = CI(i, j, k), s1 = st[dir];
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               double f = 0.0;
               for (int ph = 0; ph < 2; ph++) {
                  const double* c = ph == 0 ? cl : cs;
                  double d = ph == 0 ? DL[fi] : DA[fi];
                  if (bnd)
                     f = f + d * dinv[dir] * (c[c0] - c[c0 - s1]);
                  else
                     f = f + d * (d15 * (c[c0] - c[c0 - s1]) -
                                  d1 * (c[c0 + s1] - c[c0 - 2 * s1]));
               }
               fl[fi] = f;
            }
   }
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            double s = dinv[0] * (fl[FX(i + 1, j, k)] - fl[FX(i, j, k)]) +
                       dinv[1] * (fl[YOFF + FY(i, j + 1, k)] - fl[YOFF + FY(i, j, k)]);
            if (G->ndim == 3)
               s = s + dinv[2] * (fl[ZOFF + FZ(i, j, k + 1)] - fl[ZOFF + FZ(i, j, k)]);
            rhs[CI(i, j, k)] = p->conc_mobility * s;
         }
}

/* Multi-order phase RHS: MultiOrderRHSStrategy::evaluateRHS
 * (source/MultiOrderRHSStrategy.cc:146-259): gradient_flux of every order
 * with epsilon / sqrt(2) (source/PhaseFluxStrategyFactory.h:19-22), then
 * computerhsmultiorder (source/fortran/2d/multiorder.m4:16-105): phi2sum
 * summed over the orders in order, rhs = div(flux) - m g', g' = phi^3 - phi +
 * 2 phi gamma (phi2sum - phi), m = 8 w (:51), gamma = 1.5
 * (source/QuatModelParameters.cc:40); then rhs *= M (the mobility cell field,
 * or the uniform M when M == NULL). */
/* The side fluxes of every order: the simple gradient flux with eps / sqrt(2)
 * (source/PhaseFluxStrategyFactory.h:19-22), norder blocks of nf faces. */
void o_phase_flux_multiorder(const OGeom* G, const ampe_model_params* p, int norder,
                             const double* phi, double* fl)
{
   long ng = G->ncg;
   long nf = G->nfx + G->nfy + G->nfz;
   double eps = p->epsilon_phi / sqrt(2.);
   for (int ip = 0; ip < norder; ip++) gradient_flux(G, eps, phi + ip * ng, fl + ip * nf);
}

/* computerhsmultiorder from given side fluxes (norder blocks of nf faces). */
void o_phase_rhs_multiorder_fx(const OGeom* G, const ampe_model_params* p, int norder,
                               const double* fl, const double* phi, const double* M,
                               double* rhs)
{
   long ng = G->ncg;
   long nf = G->nfx + G->nfy + G->nfz;
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   double m = 8. * p->phi_well_scale, gamma = 1.5;
   double* s2 = (double*)malloc(sizeof(double) * ng);
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c = CI(i, j, k);
            s2[c] = phi[c] * phi[c];
         }
   for (int ip = 1; ip < norder; ip++)
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               long c = CI(i, j, k);
               s2[c] = s2[c] + phi[ip * ng + c] * phi[ip * ng + c];
            }
   for (int ip = 0; ip < norder; ip++) {
      const double* f = fl + ip * nf;
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               long c = CI(i, j, k);
               double dt = (f[FX(i + 1, j, k)] - f[FX(i, j, k)]) * dinv[0] +
                           (f[YOFF + FY(i, j + 1, k)] - f[YOFF + FY(i, j, k)]) * dinv[1];
               if (G->ndim == 3)
                  dt = dt + dinv[2] * (f[ZOFF + FZ(i, j, k + 1)] - f[ZOFF + FZ(i, j, k)]);
               double v = phi[ip * ng + c];
               double gp = v * v * v - v;
               double dsum = s2[c] - v;
               gp = gp + 2. * v * gamma * dsum;
               double r = dt - m * gp;
               rhs[ip * ng + c] = r * (M ? M[c] : p->phi_mobility);
            }
   }
   free(s2);
}

void o_phase_rhs_multiorder(const OGeom* G, const ampe_model_params* p, int norder,
                            const double* phi, const double* M, double* rhs)
{
   long nf = G->nfx + G->nfy + G->nfz;
   double* fl = (double*)malloc(sizeof(double) * nf * norder);
   o_phase_flux_multiorder(G, p, norder, phi, fl);
   o_phase_rhs_multiorder_fx(G, p, norder, fl, phi, M, rhs);
   free(fl);
}

/* projectphi (source/fortran/2d/phase.m4:51-110) over the interior cells. */
void o_project_phases(const OGeom* G, int norder, const double* phi, double* corr,
                      double* err)
{
   long ng = G->ncg;
   double avg = 1.0 / norder;
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c = CI(i, j, k);
            double fac = 0.0;
            for (int m = 0; m < norder; m++) fac = fac + phi[m * ng + c];
            fac = (fac - 1.0) * avg;
            for (int m = 0; m < norder; m++) corr[m * ng + c] = phi[m * ng + c] - fac;
            fac = 0.0;
            for (int m = 0; m < norder; m++) fac = fac + err[m * ng + c];
            fac = fac * avg;
            for (int m = 0; m < norder; m++) err[m * ng + c] = err[m * ng + c] - fac;
            for (int m = 0; m < norder; m++)
               corr[m * ng + c] = corr[m * ng + c] - phi[m * ng + c];
         }
}

/* Liquid and solid diffusivities of the face between cells cm and c0
 * (the arithmetic of o_ebs_coeffs). */
static void ebs_face_d(const ampe_model_params* p, double phm, double ph0, double* dl,
                       double* ds)
{
   double qlr = p->Q0_liquid / R_GAS, qsr = p->Q0_solid / R_GAS;
   double vphi = o_average_func(phm, ph0, p->avg_type);
   double hphi = o_interp_func(vphi, p->diff_interp);
   double invT = 2.0 / (p->temperature + p->temperature);
   *dl = (1.0 - hphi) * (p->D_liquid * exp(-qlr * invT));
   *ds = hphi * (p->D_solid * exp(-qsr * invT));
}

/* EBS flux with add_flux_iso (2d/flux.m4:69-135: the face and its two
 * transverse neighbour faces weighted 10/12 and 1/12, terms in the order
 * +t, 0, -t as (w D) dc; 3d/flux.m4:103-232: the four transverse neighbours
 * weighted 1/12 and the face 8/12, terms w (D dc) in the order x: -y, -z,
 * +y, +z; y: -x, +x, -z, +z; z: -x, +x, -y, +y; then the face), liquid then
 * solid added to a zero flux; diffusivities of every face from phid (ghosts
 * filled), then computerhsconcentration as o_conc_rhs_ebs. */
void o_conc_rhs_ebs_iso(const OGeom* G, const ampe_model_params* p, const double* phid,
                        const double* cl, const double* cs, double* rhs)
{
   long st[3] = {1, G->sy, G->sz};
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   long nf = G->nfx + G->nfy + G->nfz;
   double* fl = (double*)malloc(sizeof(double) * nf);
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
      double w1 = 1.0 / (12.0 * G->dx[dir]);
      double wc = (G->ndim == 2 ? 10.0 : 8.0) / (12.0 * G->dx[dir]);
      /* transverse offsets in the reference's order */
      long to[4];
      int nt;
      if (G->ndim == 2) {
         long t = st[1 - dir];
         to[0] = t;
         to[1] = 0;
         to[2] = -t;
         nt = 3;
      } else {
         int a = dir == 0 ? 1 : 0, b = dir == 2 ? 1 : 2;
         if (dir == 0) {
            to[0] = -st[a]; to[1] = -st[b]; to[2] = st[a]; to[3] = st[b];
         } else {
            to[0] = -st[a]; to[1] = st[a]; to[2] = -st[b]; to[3] = st[b];
         }
         nt = 4;
      }
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k);
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               double f = 0.0;
               for (int ph = 0; ph < 2; ph++) {
                  const double* cc = ph == 0 ? cl : cs;
                  double acc = f;
                  if (G->ndim == 2) {
                     for (int t = 0; t < nt; t++) {
                        long h = c0 + to[t], m = h - st[dir];
                        double dl, ds;
                        ebs_face_d(p, phid[m], phid[h], &dl, &ds);
                        double dd = ph == 0 ? dl : ds;
                        double w = to[t] == 0 ? wc : w1;
                        acc = acc + w * dd * (cc[h] - cc[m]);
                     }
                  } else {
                     for (int t = 0; t < 5; t++) {
                        long o = t < 4 ? to[t] : 0;
                        long h = c0 + o, m = h - st[dir];
                        double dl, ds;
                        ebs_face_d(p, phid[m], phid[h], &dl, &ds);
                        double dd = ph == 0 ? dl : ds;
                        double w = t < 4 ? w1 : wc;
                        acc = acc + w * (dd * (cc[h] - cc[m]));
                     }
                  }
                  f = acc;
               }
               fl[fi] = f;
            }
   }
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            double s = dinv[0] * (fl[FX(i + 1, j, k)] - fl[FX(i, j, k)]) +
                       dinv[1] * (fl[YOFF + FY(i, j + 1, k)] - fl[YOFF + FY(i, j, k)]);
            if (G->ndim == 3)
               s = s + dinv[2] * (fl[ZOFF + FZ(i, j, k + 1)] - fl[ZOFF + FZ(i, j, k)]);
            rhs[CI(i, j, k)] = p->conc_mobility * s;
         }
}

/* Grains::findAndNumberGrains (source/Grains.cc:263-645) on one patch
 * covering the domain (cell weights all positive): number the cells
 * with phase >= thr by their box index (x fastest), -1 elsewhere and in the
 * ghosts; then sweep the cells in box order, lowering each labelled cell to a
 * face neighbour's label smaller than the label it had at the start of its
 * visit (the last such neighbour in the order -x, +x, -y, +y, -z, +z wins, as
 * in the reference), with a periodic ghost fill before every sweep (ghosts
 * stay -1 on a non-periodic domain: no physical-boundary fill), until a
 * sweep changes nothing or the reference's iteration cap (twice twice the
 * largest extent) is reached; finally renumber the surviving labels 0, 1, ...
 * in increasing order.  labels: interior cells, x fastest.  Returns the
 * number of grains; *sweeps = sweeps done. */
int o_grains_label(const OGeom* G, const double* phase, double thr, int periodic, int* labels,
                   int* sweeps)
{
   long st[3] = {1, G->sy, G->sz};
   int* g = (int*)malloc(sizeof(int) * G->ncg);
   for (long c = 0; c < G->ncg; c++) g[c] = -1;
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            if (phase[c0] >= thr) g[c0] = i + G->nx * (j + G->ny * k);
         }
   int width = G->nx > G->ny ? G->nx : G->ny;
   if (G->ndim == 3 && G->nz > width) width = G->nz;
   int max_count = 2 * (2 * width);
   int counter = 0;
   while (counter < max_count) {
      int changed = 0;
      for (int k = -G->gz; k < G->nz + G->gz && periodic; k++)
         for (int j = -G->g; j < G->ny + G->g; j++)
            for (int i = -G->g; i < G->nx + G->g; i++) {
               if (i >= 0 && i < G->nx && j >= 0 && j < G->ny && k >= 0 && k < G->nz) continue;
               g[CI(i, j, k)] = g[CI(wrap(i, G->nx), wrap(j, G->ny), wrap(k, G->nz))];
            }
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               long c0 = CI(i, j, k);
               const int n = g[c0];
               if (n < 0) continue;
               for (int d = 0; d < G->ndim; d++) {
                  int nm = g[c0 - st[d]];
                  if (nm >= 0 && nm < n) {
                     g[c0] = nm;
                     changed = 1;
                  }
                  int np = g[c0 + st[d]];
                  if (np >= 0 && np < n) {
                     g[c0] = np;
                     changed = 1;
                  }
               }
            }
      counter++;
      if (!changed) break;
   }
   if (sweeps) *sweeps = counter;
   /* renumber: sorted distinct labels -> 0, 1, ... (labels are box indices) */
   long nbox = (long)G->nx * G->ny * G->nz;
   int* rank = (int*)malloc(sizeof(int) * nbox);
   for (long b = 0; b < nbox; b++) rank[b] = 0;
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            int n = g[CI(i, j, k)];
            if (n >= 0) rank[n] = 1;
         }
   int ng = 0;
   for (long b = 0; b < nbox; b++) {
      int f = rank[b];
      rank[b] = ng;
      ng += f;
   }
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            int n = g[CI(i, j, k)];
            labels[(long)i + (long)G->nx * ((long)j + (long)G->ny * k)] = n >= 0 ? rank[n] : -1;
         }
   free(rank);
   free(g);
   return ng;
}

/* tagfromgrads (2d/quattag.m4:9-55) on the cell gradient of phi
 * (QuatModel::computePhaseGradCell: diffs + cell average, quatgrad_cell
 * formula).  tags: interior cells, i fastest, no ghosts. */
void o_tag_cells(const OGeom* G, const ampe_model_params* p,
                 const double* phi, double tol, int* tags)
{
   (void)p;
   long st[3] = {1, G->sy, G->sz};
   double p5[3] = {0.5 / G->dx[0], 0.5 / G->dx[1], 0.5 / G->dx[2]};
   double tol2 = tol * tol;
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            double g = 0.0;
            for (int d = 0; d < G->ndim; d++) {
               double gd = ((phi[c0 + st[d]] - phi[c0]) + (phi[c0] - phi[c0 - st[d]])) * p5[d];
               double gh = gd * G->dx[d];
               g = g + gh * gh;
            }
            tags[(long)i + (long)G->nx * ((long)j + (long)G->ny * k)] = (g >= tol2) ? 1 : 0;
         }
}

/* ------------------------------------------------------------------------
 * Elliptic FAC operator pieces.
 * ---------------------------------------------------------------------- */

/* PhaseFACOps::setCOnPatchPrivate (source/PhaseFACOps.cc:101-183) */
void o_phase_fac_C(const OGeom* G, const ampe_model_params* p, double gamma,
                   const double* phi, const double* M, double* C)
{
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            double g2 = o_second_deriv_well_func(phi[c0], p->phi_well_type);
            double gm = gamma * (M ? M[c0] : p->phi_mobility);
            C[c0] = 1.0 + gm * p->phi_well_scale * g2;
         }
}

/* flux = dc * (u_i - u_{i-1}) / h on every face: efo_compfluxvardc2d
 * (2d/ellipticfacops.m4:16-57) and SAMRAI compfluxcondc (constant D). */
static void fac_flux(const OGeom* G, const ampe_elliptic_op* op,
                     const double* u, double* flux)
{
   long st[3] = {1, G->sy, G->sz};
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   for (int dir = 0; dir < G->ndim; dir++) {
      int ex = G->nx + (dir == 0), ey = G->ny + (dir == 1), ez = G->nz + (dir == 2);
      long off = (dir == 0 ? 0 : dir == 1 ? YOFF : ZOFF);
#pragma omp parallel for
      for (int k = 0; k < ez; k++)
         for (int j = 0; j < ey; j++)
            for (int i = 0; i < ex; i++) {
               long c0 = CI(i, j, k), cm = c0 - st[dir];
               long fi = off + (long)i + (long)ex * ((long)j + (long)ey * k);
               double d = op->D ? op->D[fi] : op->D_const;
               flux[fi] = dinv[dir] * d * (u[c0] - u[cm]);
            }
   }
}

static double op_div(const OGeom* G, const double* flux, const double* dinv,
                     int i, int j, int k)
{
   double s = dinv[0] * (flux[FX(i + 1, j, k)] - flux[FX(i, j, k)]) +
              dinv[1] * (flux[YOFF + FY(i, j + 1, k)] - flux[YOFF + FY(i, j, k)]);
   if (G->ndim == 3)
      s = s + dinv[2] * (flux[ZOFF + FZ(i, j, k + 1)] - flux[ZOFF + FZ(i, j, k)]);
   return s;
}

/* Red (color 0) or black (color 1) update: efo_rbgswithfluxmax* (2-D
 * colouring by (jfirst+j) parity, 2d/ellipticfacops.m4:111-130; 3-D by
 * (ifirst+j+k) parity, 3d/ellipticfacops.m4).  Returns max |residual|. */
double o_rbgs_color(const OGeom* G, const ampe_elliptic_op* op,
                    const double* rhs, double* u, int color, double* flux)
{
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   double maxres = 0.0;
   fac_flux(G, op, u, flux);
#pragma omp parallel for reduction(max : maxres)
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++) {
         int x = (G->ndim == 3) ? (0 + j + k) : (0 + j);
         int io = ((x - (x / 2) * 2) != color) ? 1 : 0;
         for (int i = io; i < G->nx; i += 2) {
            long c0 = CI(i, j, k);
            double m = op->M ? op->M[c0] : op->M_const;
            double cc = op->C ? op->C[c0] : op->C_const;
            double res = rhs[c0] - m * op_div(G, flux, dinv, i, j, k) - cc * u[c0];
            double dsum;
            if (op->D) {
               dsum = dinv[0] * dinv[0] * (op->D[FX(i + 1, j, k)] + op->D[FX(i, j, k)]) +
                      dinv[1] * dinv[1] *
                          (op->D[YOFF + FY(i, j + 1, k)] + op->D[YOFF + FY(i, j, k)]);
               if (G->ndim == 3)
                  dsum = dsum + dinv[2] * dinv[2] *
                                    (op->D[ZOFF + FZ(i, j, k + 1)] + op->D[ZOFF + FZ(i, j, k)]);
            } else {
               double dc = op->D_const;
               dsum = dinv[0] * dinv[0] * (dc + dc) + dinv[1] * dinv[1] * (dc + dc);
               if (G->ndim == 3) dsum = dsum + dinv[2] * dinv[2] * (dc + dc);
            }
            double dudr = 1.0 / (m * dsum - cc);
            double du = -res * dudr;
            u[c0] = u[c0] + du * 1.0;
            if (maxres < fabs(res)) maxres = fabs(res);
         }
      }
   return maxres;
}

/* EllipticFACOps::smoothErrorByRedBlack (source/EllipticFACOps.cc:1067-1230) */
double o_fac_smooth(const OGeom* G, const ampe_elliptic_op* op,
                    const double* rhs, double* u, int sweeps, double tol,
                    double* flux)
{
   double maxres = 0.0;
   int not_converged = 1;
   o_ghost_fill(G, u, 1);
   for (int s = 0; s < sweeps && not_converged; s++) {
      o_ghost_fill(G, u, 1);
      double r = o_rbgs_color(G, op, rhs, u, 0, flux);
      o_ghost_fill(G, u, 1);
      double b = o_rbgs_color(G, op, rhs, u, 1, flux);
      o_ghost_fill(G, u, 1);
      if (tol >= 0.0) {
         maxres = r > b ? r : b;
         not_converged = maxres > tol;
      } else {
         maxres = r > b ? r : b;
      }
   }
   return maxres;
}

/* efo_compres*sca (2d/ellipticfacops.m4:346-446) after a ghost fill; RMS
 * norm (EllipticFACOps::computeResidualNorm :1760-1790). */
double o_fac_residual(const OGeom* G, const ampe_elliptic_op* op,
                      const double* rhs, double* u, double* res, double* flux)
{
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   o_ghost_fill(G, u, 1);
   fac_flux(G, op, u, flux);
   double s = 0.0;
#pragma omp parallel for reduction(+ : s)
   for (int k = 0; k < G->nz; k++)
      for (int j = 0; j < G->ny; j++)
         for (int i = 0; i < G->nx; i++) {
            long c0 = CI(i, j, k);
            double m = op->M ? op->M[c0] : op->M_const;
            double cc = op->C ? op->C[c0] : op->C_const;
            double r = rhs[c0] - m * op_div(G, flux, dinv, i, j, k) - cc * u[c0];
            res[c0] = r;
            s += r * r;
         }
   return sqrt(s / ((double)G->nx * G->ny * G->nz));
}

/* ------------------------------------------------------------------------
 * Vector reductions over interior cells.  dot carries the cell volume
 * (SAMRAI control-volume weighting); wrms = sqrt(mean (x w)^2).
 * ---------------------------------------------------------------------- */
double o_dot(const OGeom* G, int ncomp, const double* x, const double* y)
{
   double cv = G->dx[0] * G->dx[1] * (G->ndim == 3 ? G->dx[2] : 1.0);
   double s = 0.0;
   for (int m = 0; m < ncomp; m++) {
#pragma omp parallel for reduction(+ : s)
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               long c0 = (long)m * G->ncg + CI(i, j, k);
               s += x[c0] * y[c0];
            }
   }
   return s * cv;
}

double o_wrms(const OGeom* G, int ncomp, const double* x, const double* w)
{
   double s = 0.0;
   for (int m = 0; m < ncomp; m++) {
#pragma omp parallel for reduction(+ : s)
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               long c0 = (long)m * G->ncg + CI(i, j, k);
               double v = x[c0] * w[c0];
               s += v * v;
            }
   }
   return sqrt(s / ((double)ncomp * G->nx * G->ny * G->nz));
}
