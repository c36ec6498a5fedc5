// Multi-order phase field (norderp > 1 without three phases):
// MultiOrderRHSStrategy::evaluateRHS (source/MultiOrderRHSStrategy.cc:146-259)
// with the Simple phase flux of epsilon / sqrt(2) (source/PhaseFluxStrategyFactory.h
// :19-22), computerhsmultiorder (source/fortran/2d/multiorder.m4:16-105,
// 3d/multiorder.m4) and the product by the phase mobility; and the projection
// onto sum(phi) = 1, projectphi (source/fortran/2d/phase.m4:51-110) as
// MultiOrderRHSStrategy::projectPhases calls it (:262-300).
//
// One thread per interior cell walks the orders: the face fluxes
// (phi_c - phi_m) e^2/dx are formed where the divergence needs them, so the
// flux field the reference materialises never reaches HBM; each order's
// components are read once for the sum of squares and once for the update.
#include <cmath>

#include "kernels.h"

namespace ampe {

namespace {

constexpr double kMultiOrderGamma = 1.5;  // QuatModelParameters d_gamma (source/QuatModelParameters.cc:40)

template <int NDIM>
__global__ void __launch_bounds__(256) k_phase_rhs_multiorder(LevelArgs a, int norder,
                                                              long long ncg, double e2dx0,
                                                              double e2dx1, double e2dx2,
                                                              double dxi0, double dxi1,
                                                              double dxi2, double m,
                                                              double mob_const,
                                                              const double* __restrict__ phi,
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
         const double t = (fhi - flo) * dxi[d];
         diff = d == 0 ? t : diff + t;
      }
      double gp = v * v * v - v;
      const double dsum = s2 - v;  // the reference subtracts phi, not phi^2
      gp = gp + 2.0 * v * kMultiOrderGamma * dsum;
      double r = diff - m * gp;
      rhs[c0 + ip * ncg] = r * mb;
   }
}

__global__ void __launch_bounds__(256) k_project_phases(LevelArgs a, int norder, long long ncg,
                                                        const double* __restrict__ phi,
                                                        double* __restrict__ corr,
                                                        double* __restrict__ err)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   const long long c0 = c.c0;
   const double avg = 1.0 / norder;
   double fac = 0.0;
   for (int ip = 0; ip < norder; ip++) fac = fac + phi[c0 + ip * ncg];
   fac = (fac - 1.0) * avg;
   for (int ip = 0; ip < norder; ip++) corr[c0 + ip * ncg] = phi[c0 + ip * ncg] - fac;
   fac = 0.0;
   for (int ip = 0; ip < norder; ip++) fac = fac + err[c0 + ip * ncg];
   fac = fac * avg;
   for (int ip = 0; ip < norder; ip++) err[c0 + ip * ncg] = err[c0 + ip * ncg] - fac;
   for (int ip = 0; ip < norder; ip++)
      corr[c0 + ip * ncg] = corr[c0 + ip * ncg] - phi[c0 + ip * ncg];
}

}  // namespace

int launch_phase_rhs_multiorder(const Level& L, const ampe_model_params& p, int norder,
                                const double* phi, const double* mobility, double* rhs,
                                hipStream_t s)
{
   if (norder < 2) {
      set_error("phase_rhs_multiorder: needs norder >= 2");
      return -1;
   }
   if (p.flux_type != AMPE_FLUX_SIMPLE) {
      set_error("phase_rhs_multiorder: only the simple phase flux is supported");
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
   // MultiOrderRHSStrategy's d_m = 8 phase_well_scale (source/MultiOrderRHSStrategy.cc:51)
   const double m = 8. * p.phi_well_scale;
   LevelArgs a = level_args(L);
   dim3 g(L.ncell_tiles), blk(TX, TY);
   if (L.ndim == 3)
      hipLaunchKernelGGL(k_phase_rhs_multiorder<3>, g, blk, 0, s, a, norder, L.ncg, e2dx[0],
                         e2dx[1], e2dx[2], dxi[0], dxi[1], dxi[2], m, p.phi_mobility, phi,
                         mobility, rhs);
   else
      hipLaunchKernelGGL(k_phase_rhs_multiorder<2>, g, blk, 0, s, a, norder, L.ncg, e2dx[0],
                         e2dx[1], e2dx[2], dxi[0], dxi[1], dxi[2], m, p.phi_mobility, phi,
                         mobility, rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_project_phases(const Level& L, int norder, const double* phi, double* corr,
                          double* err, hipStream_t s)
{
   if (norder < 1) {
      set_error("project_phases: needs norder >= 1");
      return -1;
   }
   LevelArgs a = level_args(L);
   hipLaunchKernelGGL(k_project_phases, dim3(L.ncell_tiles), dim3(TX, TY), 0, s, a, norder,
                      L.ncg, phi, corr, err);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

}  // namespace ampe
