// Point-wise physics shared by the HIP kernels: interpolation / well
// functions, face averages, the |grad q| floor, free energies and the per-cell
// KKS solve.  Operation order follows the reference Fortran so FP64 results
// agree with the CPU restatement to the last bits (transcendentals aside).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/ampe_hip.h"

#define AMPE_HD __host__ __device__ __forceinline__

#include "mathfn.h"

namespace ampe {

constexpr double kRGas = 8.314472;  // J/(K mol)

AMPE_HD double dmax(double a, double b) { return a > b ? a : b; }
AMPE_HD double dmin(double a, double b) { return a < b ? a : b; }

// interp_func (source/fortran/functions.f:17-92)
AMPE_HD double interp(double phi, char type)
{
   double t;
   switch (type) {
      case 'q': t = dmax(0.0, phi); return t * t;
      case 'p':
         t = dmax(0.0, dmin(1.0, phi));
         return t * t * t * (10.0 - 15.0 * t + 6.0 * t * t);
      case 'h':
         t = dmax(0.0, dmin(1.0, phi));
         return t * t * (3.0 - 2.0 * t);
      case 'w': t = dmax(0.0, phi); return t * t * (2.0 - t);
      case 'l':
         t = dmax(0.0, phi);
         t = dmin(1.0, t);
         return t;
      case 'm':
         t = dmax(0.0, dmin(1.0, phi));
         return t * t / (2.0 * t * (t - 1.0) + 1.0);
      case '3': t = dmax(0.0, phi); return t * t * t;
      case 's': return log(cosh(10.0 * phi)) / log(cosh(10.0));
      case 'c': return 1.0;
      default: return __builtin_nan("");
   }
}

// deriv_interp_func (source/fortran/functions.f:95-173)
AMPE_HD double dinterp(double phi, char type)
{
   double t, s;
   switch (type) {
      case 'q': t = dmax(0.0, phi); return 2.0 * t;
      case 'p':
         t = dmax(0.0, dmin(1.0, phi));
         return 30.0 * t * t * (1.0 - t) * (1.0 - t);
      case 'h':
         t = dmax(0.0, dmin(1.0, phi));
         return 6.0 * t * (1.0 - t);
      case 'w': t = dmax(0.0, phi); return t * (4.0 - 3.0 * t);
      // the reference tests phi > 0 .or. phi < 1 (functions.f:136-137), true for all phi
      case 'l': return (phi > 0.0 || phi < 1.0) ? 1.0 : 0.0;
      case 'm':
         t = dmax(0.0, dmin(1.0, phi));
         s = 2.0 * t * (t - 1.0) + 1.0;
         return 2.0 * t * (1.0 - t) / (s * s);
      case '3': t = dmax(0.0, phi); return 3.0 * t * t;
      case 's':
         t = dmax(0.0, phi);
         return 10.0 * tanh(10.0 * t) / log(cosh(10.0));
      case 'c': return 0.0;
      default: return __builtin_nan("");
   }
}

// deriv_well_func / second_deriv_well_func (source/fortran/functions.f:246-330)
AMPE_HD double dwell(double phi, char type)
{
   if (type == 'd') return 32.0 * phi * (1.0 - phi) * (1.0 - 2.0 * phi);
   return 2.0 * (phi - 1.0);
}
AMPE_HD double d2well(double phi, char type)
{
   if (type == 'd') return 32.0 * (1.0 + 6.0 * phi * (phi - 1.0));
   return 2.0;
}

// average_func (source/fortran/functions.f:333-365)
AMPE_HD double face_avg(double a, double b, char type)
{
   if (type == 'h') {
      if (a < 1.0e-16 || b < 1.0e-16) return 0.0;
      return 2.0 / (1.0 / a + 1.0 / b);
   }
   return 0.5 * (a + b);
}

// eval_grad_normi (source/fortran/quat.f:1497-1537).  x**(-0.5) is evaluated
// as 1/sqrt(x) (correctly rounded sqrt and divide: within 2 ulp of the power)
// instead of a general pow, which costs a log and an exp per face.
AMPE_HD double grad_normi(double gn2, char floor_type, double floor2,
                          double max_gni)
{
   if (floor_type == 'm') return (gn2 > floor2) ? 1.0 / sqrt(gn2) : max_gni;
   if (floor_type == 't') {
      double gng2 = gn2 * max_gni * max_gni;
      if (gng2 > 0.01) {
         double gn = sqrt(gn2);
         return tanh(max_gni * gn) / gn;
      }
      return max_gni * (1.0 - gng2 * (5.0 - 2.0 * gng2) / 15.0);
   }
   return 1.0 / sqrt(gn2 + floor2);
}

// quatmobility (source/fortran/2d/mobility.m4:13-90)
AMPE_HD double quat_mobility(const ampe_model_params& p, double ph)
{
   double qf;
   if (p.quat_mobility_type == 'e') {
      double c = p.quat_mobility_scale;
      ph = dmax(0.0, dmin(1.0, ph));
      qf = (1.0 - exp(c * ph)) / (1.0 - exp(c));
      qf = 1.0 - qf;
   } else if (p.quat_mobility_type == 'i') {
      ph = dmax(1.0e-6, dmin(1.0, ph));
      qf = dmax(0.0, (1.0 - ph) / (ph * ph));
      qf = dmin(qf, p.quat_mobility_scale);
   } else {
      ph = dmax(0.0, dmin(1.0, ph));
      qf = ph * ph * ph * (10.0 - 15.0 * ph + 6.0 * ph * ph);
      qf = 1.0 - qf;
   }
   return p.min_quat_mobility + (p.quat_mobility - p.min_quat_mobility) * qf;
}

// ---------------------------------------------------------------- energies
// Quadratic wells A (c - ceq(T))^2 and binary CALPHAD (end members, ideal
// mixing, Redlich-Kister excess), J/mol.  phase 0 = liquid, 1 = solid A.
AMPE_HD double quad_ceq(const ampe_model_params& p, int ph)
{
   return p.quad_ceq[ph] + p.quad_m[ph] * (p.temperature - p.quad_Tref);
}

// Binary CALPHAD phase at temperature T: G, dG/dc and d2G/dc2 [J/mol].
AMPE_HD double calphad_g(const ampe_calphad_phase& cp, double T, double c)
{
   double RT = kRGas * T;
   double x = 2.0 * c - 1.0;
   double P = cp.L[0] + x * (cp.L[1] + x * (cp.L[2] + x * cp.L[3]));
   return c * cp.gA + (1.0 - c) * cp.gB +
          RT * (c * log_pos(c) + (1.0 - c) * log_pos(1.0 - c)) + c * (1.0 - c) * P;
}

AMPE_HD double calphad_mu(const ampe_calphad_phase& cp, double T, double c)
{
   double RT = kRGas * T;
   double x = 2.0 * c - 1.0;
   double P = cp.L[0] + x * (cp.L[1] + x * (cp.L[2] + x * cp.L[3]));
   double dP = cp.L[1] + x * (2.0 * cp.L[2] + x * 3.0 * cp.L[3]);
   return cp.gA - cp.gB + RT * log_pos(c / (1.0 - c)) + (1.0 - 2.0 * c) * P +
          2.0 * c * (1.0 - c) * dP;
}

AMPE_HD double calphad_d2(const ampe_calphad_phase& cp, double T, double c)
{
   double RT = kRGas * T;
   double x = 2.0 * c - 1.0;
   double P = cp.L[0] + x * (cp.L[1] + x * (cp.L[2] + x * cp.L[3]));
   double dP = cp.L[1] + x * (2.0 * cp.L[2] + x * 3.0 * cp.L[3]);
   double d2P = 2.0 * cp.L[2] + x * 6.0 * cp.L[3];
   return RT / (c * (1.0 - c)) - 2.0 * P + 4.0 * (1.0 - 2.0 * c) * dP +
          4.0 * c * (1.0 - c) * d2P;
}

// Dilute binary solution (c = solute fraction):
//   f_L(c) = RT (c ln c - c),
//   f_S(c) = RT (c ln c - c) - RT ln(k_e) c + dG(T),
//   dG(T)  = RT (k_e - 1)(T - T_m) / m,
// so mu_L = mu_S gives c_S = k_e c_L and the two-phase equilibrium lies on
// the linear liquidus c_L = (T - T_m) / m.
AMPE_HD double dilute_dG(const ampe_model_params& p)
{
   double RT = kRGas * p.temperature;
   return RT * (p.dilute_keq - 1.0) * (p.temperature - p.dilute_Tm) / p.dilute_m;
}

// Temperature-dependent parts of the CALPHAD atomic mobilities, per phase
// (0 liquid, 1 solid A) and species: Q_A, Q_B and the Redlich-Kister QQ_n of
// CALPHADMobility::getQ / getQQn (source/CALPHADMobility.h:42-103), formed on
// the host once per launch.
struct MobConst {
   double RT;
   double QA[2][2], QB[2][2];
   double QQ[2][2][4];
};

inline MobConst mob_const(const ampe_model_params& p)
{
   MobConst m;
   const double T = p.temperature;
   m.RT = kRGas * T;
   for (int ph = 0; ph < 2; ph++)
      for (int sp = 0; sp < 2; sp++) {
         const ampe_calphad_mobility& mb = p.mobility[ph][sp];
         double q[2];
         for (int x = 0; x < 2; x++) {
            q[x] = mb.q[x][0] + m.RT * std::log(mb.q[x][1]);
            if (mb.has_extra) q[x] += mb.qx[x][0] * T * std::log(T) + mb.qx[x][1] / T;
         }
         m.QA[ph][sp] = q[0];
         m.QB[ph][sp] = q[1];
         for (int n = 0; n < 4; n++)
            m.QQ[ph][sp][n] = mb.qq[n][0] + m.RT * std::log(mb.qq[n][1]);
      }
   return m;
}

// CALPHADMobility::getAtomicMobility for a binary (getDeltaG,
// source/CALPHADMobility.cc:155-166): exp(dG / RT) / RT.
AMPE_HD double atomic_mobility(const MobConst& m, int ph, int sp, double c0, double c1)
{
   const double dc = c0 - c1;
   const double* qq = m.QQ[ph][sp];
   const double dG = c0 * m.QA[ph][sp] + c1 * m.QB[ph][sp] +
                     c0 * c1 * (qq[0] + dc * (qq[1] + dc * (qq[2] + dc * qq[3])));
   const double rtinv = 1.0 / m.RT;
   return exp(dG * rtinv) * rtinv;
}

// computeDiffusionMobilityBinaryPhase (source/CALPHADMobility.cc:200-218):
// c0 c1 (c0 m1 + c1 m0), in um^2 (the 1e12 of m^2).
AMPE_HD double phase_diffusion_mobility(const MobConst& m, int ph, double c0)
{
   const double c1 = 1.0 - c0;
   const double m0 = atomic_mobility(m, ph, 0, c0, c1);
   const double m1 = atomic_mobility(m, ph, 1, c0, c1);
   const double mm = c0 * m1 + c1 * m0;
   return c0 * c1 * mm * 1.0e12;
}

// The energy functions take the model as a template argument FE when a kernel
// is specialised for it; FE < 0 reads p.fe_model at run time.
template <int FE>
AMPE_HD double fenergy_t(const ampe_model_params& p, double c, int ph)
{
   const int fe = (FE < 0) ? p.fe_model : FE;
   if (fe == AMPE_FE_QUADRATIC) {
      double d = c - quad_ceq(p, ph);
      return p.quad_A[ph] * d * d;
   }
   if (fe == AMPE_FE_KKS_DILUTE) {
      double RT = kRGas * p.temperature;
      double f = RT * (c * log_pos(c) - c);
      if (ph == 1) f = f - RT * log_pos(p.dilute_keq) * c + dilute_dG(p);
      return f;
   }
   return calphad_g(p.calphad[ph], p.temperature, c);
}

template <int FE>
AMPE_HD double dfenergy_t(const ampe_model_params& p, double c, int ph)
{
   const int fe = (FE < 0) ? p.fe_model : FE;
   if (fe == AMPE_FE_QUADRATIC) return 2.0 * p.quad_A[ph] * (c - quad_ceq(p, ph));
   if (fe == AMPE_FE_KKS_DILUTE) {
      double RT = kRGas * p.temperature;
      double mu = RT * log_pos(c);
      if (ph == 1) mu = mu - RT * log_pos(p.dilute_keq);
      return mu;
   }
   return calphad_mu(p.calphad[ph], p.temperature, c);
}

template <int FE>
AMPE_HD double d2fenergy_t(const ampe_model_params& p, double c, int ph)
{
   const int fe = (FE < 0) ? p.fe_model : FE;
   if (fe == AMPE_FE_QUADRATIC) return 2.0 * p.quad_A[ph];
   if (fe == AMPE_FE_KKS_DILUTE) return kRGas * p.temperature / c;
   return calphad_d2(p.calphad[ph], p.temperature, c);
}

AMPE_HD double fenergy(const ampe_model_params& p, double c, int ph)
{
   return fenergy_t<-1>(p, c, ph);
}
AMPE_HD double dfenergy(const ampe_model_params& p, double c, int ph)
{
   return dfenergy_t<-1>(p, c, ph);
}
AMPE_HD double d2fenergy(const ampe_model_params& p, double c, int ph)
{
   return d2fenergy_t<-1>(p, c, ph);
}

// Phase mobility at one cell: the uniform M_phi, or Kim's KKS mobility
// (KimMobilityStrategyInfMob::evaluateMobility, source/
// KimMobilityStrategyInfMob.cc:54-83; FiniteMob, KimMobilityStrategyFiniteMob.h
// :29-38; FiniteMobAntiTrap, .cc:54-80) with kks_mobility_factor
// (source/KKStools.cc:13-30) and zeta = (cL - cS) f_L''(cL) (cL - cS) / D_L(T)
// converted by 1e-6 / v_m.
template <int FE>
AMPE_HD double phase_mobility_t(const ampe_model_params& p, double cl, double cs)
{
   if (p.phi_mobility_type != AMPE_PHI_MOB_KIM) return p.phi_mobility;
   double eps = p.epsilon_phi, w = p.phi_well_scale;
   double a2 = (p.energy_interp == 'p') ? 47. / 60. : 0.5;
   double factor = 16. * w / (3. * eps * eps * a2);
   double alpha = 0.0;
   bool finite = p.kim_interface_mobility > 0.0;
   if (finite) {
      double xi = eps / sqrt(16. * w);
      alpha = 3. * sqrt(2.) * xi / p.kim_interface_mobility;
      if (!p.kim_antitrapping) return 1. / alpha;
   }
   double T = p.temperature;
   double DL = p.D_liquid * exp(-p.Q0_liquid / (kRGas * T));
   double d2 = d2fenergy_t<FE>(p, cl, 0);
   double zeta = 0.;
   zeta += (cl - cs) * d2 * (cl - cs);
   double zf = zeta / DL;
   zf *= (1.e-6 / p.vm_L);
   if (finite) return 1. / (alpha + (1. / factor) * zf);
   return factor / zf;
}

AMPE_HD double phase_mobility(const ampe_model_params& p, double cl, double cs)
{
   return phase_mobility_t<-1>(p, cl, cs);
}

// End-member G(T) of a tdep table: interval k is the last with T >= Tc[k]
// (k <= nrange - 1); terms added in database order.
AMPE_HD double calphad_gfun(const ampe_calphad_gfun& f, double T)
{
   int k = 0;
   while (k < f.nrange - 1 && T >= f.Tc[k + 1]) k++;
   const double* a = f.coef[k];
   double T2 = T * T, T3 = T2 * T;
   double T7 = T3 * T3 * T;
   double Tm9 = 1.0 / (T3 * T3 * T3);
   double g = a[0] + a[1] * T + a[2] * T * log_pos(T);
   g = g + a[3] * T2;
   g = g + a[4] * T3;
   g = g + a[5] * T7;
   g = g + a[6] / T;
   g = g + a[7] * Tm9;
   return g;
}

AMPE_HD void calphad_at(const ampe_calphad_tdep& tab, double T, ampe_calphad_phase* out)
{
   for (int ph = 0; ph < 2; ph++) {
      out[ph].gA = calphad_gfun(tab.g[ph][0], T);
      out[ph].gB = calphad_gfun(tab.g[ph][1], T);
      for (int i = 0; i < 4; i++) out[ph].L[i] = tab.L[ph][i][0] + tab.L[ph][i][1] * T;
   }
}

// Residual of the binary CALPHAD KKS system at (xl, xs):
// F = [(mu_L(xl) - mu_S(xs))/RT, (1-h) xl + h xs - c].
AMPE_HD void kks_residual(const ampe_calphad_phase* ph, double T, double RTinv, double h,
                          double c, double xl, double xs, double& f1, double& f2)
{
   f1 = (calphad_mu(ph[0], T, xl) - calphad_mu(ph[1], T, xs)) * RTinv;
   f2 = (1.0 - h) * xl + h * xs - c;
}

// Start of a KKS Newton: the guess clamped into [1e-12, 1-1e-12].
AMPE_HD void kks_clamp(double& xl, double& xs)
{
   const double lo = 1.0e-12, hi = 1.0 - 1.0e-12;
   if (!(xl > lo)) xl = lo;
   if (!(xl < hi)) xl = hi;
   if (!(xs > lo)) xs = lo;
   if (!(xs < hi)) xs = hi;
}

// Newton iterations from a clamped start (xl, xs) whose residual (f1, f2) is
// given; step halving keeps the iterate inside (0,1); converged when both
// |F_i| < tol.  Returns the number of updates, -1 after maxits updates
// without convergence, or -2 when probe_its >= 0 and max|F| > probe_tol after
// probe_its updates (the caller restarts).
AMPE_HD int kks_newton_from(const ampe_calphad_phase* ph, double T, double h, double c,
                            double& cl, double& cs, double f1, double f2, int maxits,
                            double tol, int probe_its, double probe_tol)
{
   double RTinv = 1.0 / (kRGas * T);
   double xl = cl, xs = cs;
   for (int it = 0; it <= maxits; it++) {
      if (it > 0) kks_residual(ph, T, RTinv, h, c, xl, xs, f1, f2);
      if (fabs(f1) < tol && fabs(f2) < tol) {
         cl = xl;
         cs = xs;
         return it;
      }
      if (it == maxits) break;
      if (it == probe_its && !(dmax(fabs(f1), fabs(f2)) <= probe_tol)) {
         cl = xl;
         cs = xs;
         return -2;
      }
      double a = calphad_d2(ph[0], T, xl) * RTinv;
      double b = -calphad_d2(ph[1], T, xs) * RTinv;
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
   cl = xl;
   cs = xs;
   return -1;
}

// Binary CALPHAD KKS Newton from the guess (cl, cs) (see kks_newton_from).
AMPE_HD int kks_newton(const ampe_calphad_phase* ph, double T, double h, double c, double& cl,
                       double& cs, int maxits, double tol, int probe_its, double probe_tol)
{
   kks_clamp(cl, cs);
   double f1, f2;
   kks_residual(ph, T, 1.0 / (kRGas * T), h, c, cl, cs, f1, f2);
   return kks_newton_from(ph, T, h, c, cl, cs, f1, f2, maxits, tol, probe_its, probe_tol);
}

// Quadratic wells: closed form of mu_L = mu_S with mass balance.
AMPE_HD void kks_quadratic(const ampe_model_params& p, double h, double c, double& cl,
                           double& cs)
{
   double eL = quad_ceq(p, 0), eS = quad_ceq(p, 1);
   double r = p.quad_A[0] / p.quad_A[1];
   double xl = (c - h * (eS - r * eL)) / (1.0 - h + h * r);
   cl = xl;
   cs = eS + r * (xl - eL);
}

// Per-cell KKS at the model temperature: mu_L(cL) = mu_S(cS),
// (1-h) cL + h cS = c.  Quadratic wells and the dilute solution: closed
// forms.  CALPHAD: kks_newton from the guess in (cl, cs).  Returns the
// iteration count or -1.
AMPE_HD int kks_cell(const ampe_model_params& p, double c, double phi, double& cl,
                     double& cs)
{
   double h = interp(phi, p.conc_interp);
   if (p.fe_model == AMPE_FE_QUADRATIC) {
      kks_quadratic(p, h, c, cl, cs);
      return 0;
   }
   if (p.fe_model == AMPE_FE_KKS_DILUTE) {
      double k = p.dilute_keq;
      double xl = c / (1.0 - h + h * k);
      cl = xl;
      cs = k * xl;
      return 0;
   }
   return kks_newton(p.calphad, p.temperature, h, c, cl, cs, p.newton_max_its, p.newton_tol,
                     -1, 0.0);
}

}  // namespace ampe
