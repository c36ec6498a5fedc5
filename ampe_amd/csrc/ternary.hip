// Three phases (liquid, solid A, solid B) on norderp > 1 order parameters
// with a ternary quadratic model (fe_model = AMPE_FE_QUADRATIC_TERNARY_3PH):
//  - phase concentrations in closed form (the KKS conditions of the
//    QuadraticEquilibriumThreePhasesTernaryStrategyMultiOrder strategy,
//    source/QuadraticEquilibriumThreePhasesTernaryStrategyMultiOrder.cc:49-181,
//    for f_a(c) = sum_k A_ak (c_k - e_ak)^2: one diffusion potential
//    mu_k = 2 A_ak (x_ak - e_ak) in every phase and c_k = sum_a g(h_a) x_ak,
//    so mu_k = (c_k - sum_a g_a e_ak) / sum_a (g_a / 2 A_ak));
//  - the multi-order phase RHS (MultiOrderRHSStrategy::evaluateRHS,
//    source/MultiOrderRHSStrategy.cc:146-259) with Moelans' driving force of
//    QuadraticFreeEnergyStrategyMultiOrderTernaryThreePhase::addDrivingForce
//    (source/QuadraticFreeEnergyStrategyMultiOrderTernaryThreePhase.cc:237-520);
//  - face diffusivities of the three phases
//    (concentration_pfmdiffusion_of_temperature_threephases, source/fortran/
//    2d/concentrationdiffusion.m4:523-640) and the EBS composition RHS of two
//    species (EBSCompositionRHSStrategy::computeFluxOnPatch with add_flux for
//    L, A then B, source/EBSCompositionRHSStrategy.cc:183-328).
// Phase weights: h = (phi_L^2, sum over A's orders phi^2, sum over B's) /
// sum phi^2, the liquid being the last order.  The six phase concentrations
// are stored as fields cL0, cL1, cA0, cA1, cB0, cB1.
#include <cmath>

#include "kernels.h"
#include "physics.h"

namespace ampe {

namespace {

constexpr double kMultiOrderGammaT = 1.5;  // QuatModelParameters d_gamma (source/QuatModelParameters.cc:40)

struct Tern3 {
   double A[3][2], e[3][2];
   double conv[3];  // 1e-6 / v_m of L, A, B
   int nA;          // orders of solid A
   char conc_interp;
};

Tern3 tern3_of(const ampe_model_params& p)
{
   Tern3 t;
   for (int a = 0; a < 3; a++)
      for (int k = 0; k < 2; k++) {
         t.A[a][k] = p.tern_A[a][k];
         t.e[a][k] = p.tern_ceq[a][k];
      }
   t.conv[0] = 1.e-6 / p.vm_L;
   t.conv[1] = 1.e-6 / p.vm_A;
   t.conv[2] = 1.e-6 / p.vm_B;
   t.nA = p.norderp_A;
   t.conc_interp = p.conc_interp;
   return t;
}

// phase weights of cell c; returns 1 / sum phi^2
__device__ __forceinline__ double tern3_weights(const Tern3& t, int norder,
                                                const double* __restrict__ phi, long long ncg,
                                                long long c, double h[3])
{
   const double vl = phi[(long long)(norder - 1) * ncg + c];
   double hl = vl * vl;
   double ha = 0.0, hb = 0.0;
   for (int i = 0; i < t.nA; i++) {
      const double v = phi[(long long)i * ncg + c];
      ha += v * v;
   }
   for (int i = t.nA; i < norder - 1; i++) {
      const double v = phi[(long long)i * ncg + c];
      hb += v * v;
   }
   const double sum2 = hl + ha + hb;
   const double inv = 1. / sum2;
   h[0] = hl * inv;
   h[1] = ha * inv;
   h[2] = hb * inv;
   return inv;
}

__global__ void __launch_bounds__(256) k_kks_tern3(long long ncg, Tern3 t, int norder,
                                                   const double* __restrict__ phi,
                                                   const double* __restrict__ c,
                                                   double* __restrict__ ph)
{
   long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (n >= ncg) return;
   double h[3];
   tern3_weights(t, norder, phi, ncg, n, h);
   double g[3];
   for (int a = 0; a < 3; a++) g[a] = interp(h[a], t.conc_interp);
   for (int k = 0; k < 2; k++) {
      double num = c[(long long)k * ncg + n], den = 0.0;
      for (int a = 0; a < 3; a++) {
         num = num - g[a] * t.e[a][k];
         den = den + g[a] / (2.0 * t.A[a][k]);
      }
      const double mu = num / den;
      for (int a = 0; a < 3; a++)
         ph[(long long)(2 * a + k) * ncg + n] = t.e[a][k] + mu / (2.0 * t.A[a][k]);
   }
}

// free energy [pJ/um^3] and potentials of phase a at x
__device__ __forceinline__ double tern3_energy(const Tern3& t, int a, const double x[2],
                                               double mu[2])
{
   double f = 0.0;
   for (int k = 0; k < 2; k++) {
      const double d = x[k] - t.e[a][k];
      f = f + t.A[a][k] * d * d;
      mu[k] = 2.0 * t.A[a][k] * d * t.conv[a];
   }
   return f * t.conv[a];
}

template <int NDIM>
__global__ void __launch_bounds__(256) k_phase_rhs_tern3(LevelArgs a, Tern3 t, int norder,
                                                         long long ncg, double e2dx0,
                                                         double e2dx1, double e2dx2,
                                                         double dxi0, double dxi1, double dxi2,
                                                         double m, double mob_const,
                                                         const double* __restrict__ phi,
                                                         const double* __restrict__ ph,
                                                         const double* __restrict__ mob,
                                                         double* __restrict__ rhs)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   const long long c0 = c.c0;
   const double e2dx[3] = {e2dx0, e2dx1, e2dx2};
   const double dxi[3] = {dxi0, dxi1, dxi2};
   double s2 = phi[c0] * phi[c0];
   for (int ip = 1; ip < norder; ip++) {
      const double v = phi[c0 + ip * ncg];
      s2 = s2 + v * v;
   }
   // driving-force factors of the cell (Moelans, Acta Mater. 59 (2011))
   const double cl[2] = {ph[c0], ph[ncg + c0]};
   const double ca[2] = {ph[2 * ncg + c0], ph[3 * ncg + c0]};
   const double cb[2] = {ph[4 * ncg + c0], ph[5 * ncg + c0]};
   double muL[2], muA[2], muB[2];
   const double fl = tern3_energy(t, 0, cl, muL);
   const double fa = tern3_energy(t, 1, ca, muA);
   const double fb = tern3_energy(t, 2, cb, muB);
   const double dfa = (fa - muA[0] * ca[0] - muA[1] * ca[1]);
   const double dfb = (fb - muB[0] * cb[0] - muB[1] * cb[1]);
   const double dfl = (fl - muL[0] * cl[0] - muL[1] * cl[1]);
   double h[3];
   const double sum2inv = tern3_weights(t, norder, phi, ncg, c0, h);
   const double gA = (1. - h[1]) * dfa - h[0] * dfl - h[2] * dfb;
   const double gB = (1. - h[2]) * dfb - h[0] * dfl - h[1] * dfa;
   const double gL = (1. - h[0]) * dfl - h[1] * dfa - h[2] * dfb;
   const double mb = mob ? mob[c0] : mob_const;
   for (int ip = 0; ip < norder; ip++) {
      const double* f = phi + ip * ncg;
      const double v = f[c0];
      double diff = 0.0;
#pragma unroll
      for (int d = 0; d < NDIM; d++) {
         const long long st = c.st[d];
         const double flo = (v - f[c0 - st]) * e2dx[d];
         const double fhi = (f[c0 + st] - v) * e2dx[d];
         const double tt = (fhi - flo) * dxi[d];
         diff = d == 0 ? tt : diff + tt;
      }
      double gp = v * v * v - v;
      const double dsum = s2 - v;  // the reference subtracts phi, not phi^2
      gp = gp + 2.0 * v * kMultiOrderGammaT * dsum;
      double r = diff - m * gp;
      const double g = ip < t.nA ? gA : (ip < norder - 1 ? gB : gL);
      r = r - 2. * v * g * sum2inv;
      rhs[c0 + ip * ncg] = r * mb;
   }
}

struct Ebs3Const {
   double dl, da, db;  // D0 exp(-Q/RT) of L, A, B
   int nA, norder;
   char avg, interp;
};

template <int NDIM>
__global__ void __launch_bounds__(256) k_ebs3_coeffs(LevelArgs a, Ebs3Const k, long long ncg,
                                                     const double* __restrict__ phi,
                                                     double* __restrict__ DL,
                                                     double* __restrict__ DA,
                                                     double* __restrict__ DB)
{
   CellPos c;
   if (!cell_pos(a, 1, c)) return;
   if (NDIM == 2 && c.k > 0) return;
   for (int d = 0; d < NDIM; d++) {
      bool ok = (c.i < c.pd.nx + (d == 0)) && (c.j < c.pd.ny + (d == 1)) &&
                (NDIM == 2 || c.k < c.pd.nz + (d == 2));
      if (!ok) continue;
      const long long c0 = c.c0, cm = c0 - c.st[d];
      const long long fi = face_idx(a, c.pd, d, c.i, c.j, c.k);
      double vphi = 0.0;
      for (int ip = 0; ip < k.nA; ip++)
         vphi = vphi + face_avg(phi[ip * ncg + cm], phi[ip * ncg + c0], k.avg);
      const double vA = interp(vphi, k.interp);
      vphi = 0.0;
      for (int ip = k.nA; ip < k.norder - 1; ip++)
         vphi = vphi + face_avg(phi[ip * ncg + cm], phi[ip * ncg + c0], k.avg);
      const double vB = interp(vphi, k.interp);
      const long long l = (long long)(k.norder - 1) * ncg;
      vphi = face_avg(phi[l + cm], phi[l + c0], k.avg);
      const double vL = interp(vphi, k.interp);
      DL[fi] = vL * k.dl;
      DA[fi] = vA * k.da;
      DB[fi] = vB * k.db;
   }
}

// two species: flux_k = (0 + dxinv D_L dcL_k + dxinv D_A dcA_k) + dxinv D_B dcB_k
template <int NDIM>
__global__ void __launch_bounds__(256) k_conc_rhs_ebs3(LevelArgs a, double dxi0, double dxi1,
                                                       double dxi2, double mc, long long ncg,
                                                       const double* __restrict__ ph,
                                                       const double* __restrict__ DL,
                                                       const double* __restrict__ DA,
                                                       const double* __restrict__ DB,
                                                       double* __restrict__ rhs)
{
   AMPE_COLUMN_BEGIN(a, c)
   const long long c0 = c.c0;
   const int ijk[3] = {c.i, c.j, c.k};
   const double dxi[3] = {dxi0, dxi1, dxi2};
   for (int sp = 0; sp < 2; sp++) {
      const double* cl = ph + (long long)sp * ncg;
      const double* ca = ph + (long long)(2 + sp) * ncg;
      const double* cb = ph + (long long)(4 + sp) * ncg;
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < NDIM; d++) {
         const double dinv = dxi[d];
         int lo[3] = {ijk[0], ijk[1], ijk[2]};
         const long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
         lo[d] += 1;
         const long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
         const long long cm = c0 - c.st[d], cp = c0 + c.st[d];
         double jl = 0.0;
         jl = jl + dinv * DL[flo] * (cl[c0] - cl[cm]);
         jl = jl + dinv * DA[flo] * (ca[c0] - ca[cm]);
         jl = jl + dinv * DB[flo] * (cb[c0] - cb[cm]);
         double jh = 0.0;
         jh = jh + dinv * DL[fhi] * (cl[cp] - cl[c0]);
         jh = jh + dinv * DA[fhi] * (ca[cp] - ca[c0]);
         jh = jh + dinv * DB[fhi] * (cb[cp] - cb[c0]);
         const double tt = dinv * (jh - jl);
         s = (d == 0) ? tt : s + tt;
      }
      rhs[(long long)sp * ncg + c0] = mc * s;
   }
   AMPE_COLUMN_END
}

// Composite fluxes on a patch hierarchy: the coarse level's fused kernel
// used the coarse face flux on each coarse-fine face; one thread per listed
// coarse cell replaces it in the divergence with the area-weighted average
// of the finer level's fluxes (the side-data CONSERVATIVE_COARSEN of
// MultiOrderRHSStrategy::evaluateRHS, source/MultiOrderRHSStrategy.cc:85-138,
// and QuatIntegrator::evaluateConcentrationRHS, source/QuatIntegrator.cc:
// 2610-2620), every order and both species.
template <int NDIM>
__global__ void __launch_bounds__(256) k_hier_tern3_phase_corr(
    const CfCorr* __restrict__ cc, int n, const CfFace* __restrict__ ff, int norder,
    long long ncg, double e2c0, double e2c1, double e2c2, double e2f0, double e2f1, double e2f2,
    double dxi0, double dxi1, double dxi2, double mob, double af, double ac,
    const double* __restrict__ phi, double* __restrict__ rhs)
{
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfCorr& e = cc[t];
   const int nf = NDIM == 3 ? 4 : 2;
   const double e2c[3] = {e2c0, e2c1, e2c2}, e2f[3] = {e2f0, e2f1, e2f2};
   const double dxi[3] = {dxi0, dxi1, dxi2};
   for (int ip = 0; ip < norder; ip++) {
      const double* f = phi + (long long)ip * ncg;
      double delta = 0.0;
      for (int k = 0; k < e.nface; k++) {
         const int d = e.d[k];
         const double fold = (f[e.c0[k]] - f[e.cm[k]]) * e2c[d];
         double acc = 0.0;
         for (int r = 0; r < nf; r++) {
            const CfFace& g = ff[e.fine[k] + r];
            acc = acc + (f[g.c0] - f[g.cm]) * e2f[d] * af;
         }
         const double t1 = (acc / ac - fold) * dxi[d];
         delta = delta + (e.side[k] ? t1 : -t1);
      }
      rhs[(long long)ip * ncg + e.cell] = rhs[(long long)ip * ncg + e.cell] + mob * delta;
   }
}

__device__ __forceinline__ double ebs3_face_flux(const double* __restrict__ cl,
                                                 const double* __restrict__ ca,
                                                 const double* __restrict__ cb, double dinv,
                                                 double dl, double da, double db, long long c0,
                                                 long long cm)
{
   double j = 0.0;
   j = j + dinv * dl * (cl[c0] - cl[cm]);
   j = j + dinv * da * (ca[c0] - ca[cm]);
   j = j + dinv * db * (cb[c0] - cb[cm]);
   return j;
}

template <int NDIM>
__global__ void __launch_bounds__(256) k_hier_ebs3_corr(
    const CfCorr* __restrict__ cc, int n, const CfFace* __restrict__ ff, long long ncg,
    double dxc0, double dxc1, double dxc2, double dxf0, double dxf1, double dxf2, double mc,
    double af, double ac, const double* __restrict__ ph, const double* __restrict__ DL,
    const double* __restrict__ DA, const double* __restrict__ DB, double* __restrict__ rhs)
{
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfCorr& e = cc[t];
   const int nf = NDIM == 3 ? 4 : 2;
   const double dic[3] = {dxc0, dxc1, dxc2}, dif[3] = {dxf0, dxf1, dxf2};
   for (int sp = 0; sp < 2; sp++) {
      const double* cl = ph + (long long)sp * ncg;
      const double* ca = ph + (long long)(2 + sp) * ncg;
      const double* cb = ph + (long long)(4 + sp) * ncg;
      double delta = 0.0;
      for (int k = 0; k < e.nface; k++) {
         const int d = e.d[k];
         const long long fo = e.face[k];
         const double fold =
             ebs3_face_flux(cl, ca, cb, dic[d], DL[fo], DA[fo], DB[fo], e.c0[k], e.cm[k]);
         double acc = 0.0;
         for (int r = 0; r < nf; r++) {
            const CfFace& g = ff[e.fine[k] + r];
            acc = acc + ebs3_face_flux(cl, ca, cb, dif[d], DL[g.face], DA[g.face], DB[g.face],
                                       g.c0, g.cm) *
                            af;
         }
         const double t1 = dic[d] * (acc / ac - fold);
         delta = delta + (e.side[k] ? t1 : -t1);
      }
      rhs[(long long)sp * ncg + e.cell] = rhs[(long long)sp * ncg + e.cell] + mc * delta;
   }
}

}  // namespace

static int tern3_check(const ampe_model_params& p, int norder)
{
   if (p.fe_model != AMPE_FE_QUADRATIC_TERNARY_3PH || norder < 3 || p.norderp_A < 1 ||
       p.norderp_A > norder - 1) {
      set_error("ternary three-phase: needs fe_model AMPE_FE_QUADRATIC_TERNARY_3PH, "
                "norderp >= 3 and 1 <= norderp_A <= norderp - 1");
      return -1;
   }
   return 0;
}

int launch_kks_tern3(const Level& L, const ampe_model_params& p, int norder, const double* phi,
                     const double* c, double* ph, hipStream_t s)
{
   if (tern3_check(p, norder)) return -1;
   hipLaunchKernelGGL(k_kks_tern3, dim3((unsigned)((L.ncg + 255) / 256)), dim3(256), 0, s,
                      L.ncg, tern3_of(p), norder, phi, c, ph);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_phase_rhs_tern3(const Level& L, const ampe_model_params& p, int norder,
                           const double* phi, const double* ph, const double* mobility,
                           double* rhs, hipStream_t s)
{
   if (tern3_check(p, norder)) return -1;
   if (p.flux_type != AMPE_FLUX_SIMPLE) {
      set_error("phase_rhs_tern3: only the simple phase flux is supported");
      return -1;
   }
   const double eps = p.epsilon_phi / std::sqrt(2.);
   const double e2 = eps * eps;
   double e2dx[3], dxi[3];
   for (int d = 0; d < 3; d++) {
      const double h = d < L.ndim ? p.dx[d] : 1.0;
      e2dx[d] = e2 / h;
      dxi[d] = 1.0 / h;
   }
   const double m = 8. * p.phi_well_scale;
   LevelArgs a = level_args(L);
   dim3 g(L.ncell_tiles), blk(TX, TY);
   Tern3 t = tern3_of(p);
   if (L.ndim == 3)
      hipLaunchKernelGGL(k_phase_rhs_tern3<3>, g, blk, 0, s, a, t, norder, L.ncg, e2dx[0],
                         e2dx[1], e2dx[2], dxi[0], dxi[1], dxi[2], m, p.phi_mobility, phi, ph,
                         mobility, rhs);
   else
      hipLaunchKernelGGL(k_phase_rhs_tern3<2>, g, blk, 0, s, a, t, norder, L.ncg, e2dx[0],
                         e2dx[1], e2dx[2], dxi[0], dxi[1], dxi[2], m, p.phi_mobility, phi, ph,
                         mobility, rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_ebs3_coeffs(const Level& L, const ampe_model_params& p, int norder, const double* phi,
                       double* DL, double* DA, double* DB, hipStream_t s)
{
   if (tern3_check(p, norder)) return -1;
   Ebs3Const k;
   const double invT = 2.0 / (p.temperature + p.temperature);
   k.dl = p.D_liquid * std::exp(-(p.Q0_liquid / kRGas) * invT);
   k.da = p.D_solid * std::exp(-(p.Q0_solid / kRGas) * invT);
   k.db = p.D_solid_B * std::exp(-(p.Q0_solid_B / kRGas) * invT);
   k.nA = p.norderp_A;
   k.norder = norder;
   k.avg = p.avg_type;
   k.interp = p.diff_interp;
   LevelArgs a = level_args(L, true);
   dim3 g(L.nface_tiles), blk(TX, TY);
   if (L.ndim == 3)
      hipLaunchKernelGGL(k_ebs3_coeffs<3>, g, blk, 0, s, a, k, L.ncg, phi, DL, DA, DB);
   else
      hipLaunchKernelGGL(k_ebs3_coeffs<2>, g, blk, 0, s, a, k, L.ncg, phi, DL, DA, DB);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_conc_rhs_ebs3(const Level& L, const ampe_model_params& p, const double* ph,
                         const double* DL, const double* DA, const double* DB, double* rhs,
                         hipStream_t s)
{
   double dxi[3];
   for (int d = 0; d < 3; d++) dxi[d] = 1.0 / p.dx[d];
   LevelArgs a = level_args_col(L);
   dim3 g(L.ncol_tiles), blk(TX, TY);
   if (L.ndim == 3)
      hipLaunchKernelGGL(k_conc_rhs_ebs3<3>, g, blk, 0, s, a, dxi[0], dxi[1], dxi[2],
                         p.conc_mobility, L.ncg, ph, DL, DA, DB, rhs);
   else
      hipLaunchKernelGGL(k_conc_rhs_ebs3<2>, g, blk, 0, s, a, dxi[0], dxi[1], dxi[2],
                         p.conc_mobility, L.ncg, ph, DL, DA, DB, rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

static void tern3_e2dx(const ampe_model_params& p, int ndim, double e2dx[3], double dxi[3])
{
   const double eps = p.epsilon_phi / std::sqrt(2.);
   const double e2 = eps * eps;
   for (int d = 0; d < 3; d++) {
      const double h = d < ndim ? p.dx[d] : 1.0;
      e2dx[d] = e2 / h;
      dxi[d] = 1.0 / h;
   }
}

int launch_hier_tern3_phase_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                                 const ampe_model_params& pc, const ampe_model_params& pf,
                                 int norder, long long ncg, const double* phi, double af,
                                 double ac, double* rhs, hipStream_t s)
{
   if (ncorr <= 0) return 0;
   double ec[3], ef[3], dc[3], df[3];
   tern3_e2dx(pc, ndim, ec, dc);
   tern3_e2dx(pf, ndim, ef, df);
   dim3 g((ncorr + 255) / 256), b(256);
   if (ndim == 3)
      hipLaunchKernelGGL(k_hier_tern3_phase_corr<3>, g, b, 0, s, cc, ncorr, ff, norder, ncg,
                         ec[0], ec[1], ec[2], ef[0], ef[1], ef[2], dc[0], dc[1], dc[2],
                         pc.phi_mobility, af, ac, phi, rhs);
   else
      hipLaunchKernelGGL(k_hier_tern3_phase_corr<2>, g, b, 0, s, cc, ncorr, ff, norder, ncg,
                         ec[0], ec[1], ec[2], ef[0], ef[1], ef[2], dc[0], dc[1], dc[2],
                         pc.phi_mobility, af, ac, phi, rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_hier_ebs3_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                          const ampe_model_params& pc, const ampe_model_params& pf, long long ncg,
                          const double* ph, const double* DL, const double* DA, const double* DB,
                          double af, double ac, double* rhs, hipStream_t s)
{
   if (ncorr <= 0) return 0;
   double dc[3], df[3];
   for (int d = 0; d < 3; d++) {
      dc[d] = 1.0 / pc.dx[d];
      df[d] = 1.0 / pf.dx[d];
   }
   dim3 g((ncorr + 255) / 256), b(256);
   if (ndim == 3)
      hipLaunchKernelGGL(k_hier_ebs3_corr<3>, g, b, 0, s, cc, ncorr, ff, ncg, dc[0], dc[1],
                         dc[2], df[0], df[1], df[2], pc.conc_mobility, af, ac, ph, DL, DA, DB,
                         rhs);
   else
      hipLaunchKernelGGL(k_hier_ebs3_corr<2>, g, b, 0, s, cc, ncorr, ff, ncg, dc[0], dc[1],
                         dc[2], df[0], df[1], df[2], pc.conc_mobility, af, ac, ph, DL, DA, DB,
                         rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

}  // namespace ampe
