// Orientation-field transfers between a level and the level with every patch
// box halved (ratio 2): QuatWeightedAverage (coarsen, source/
// QuatWeightedAverage.cc:100-171, F77 quatcoarsen, source/fortran/2d/
// quatcoarsen.m4, 3d/quatcoarsen.m4) and QuatLinearRefine (refine, source/
// QuatLinearRefine.cc:107-174, F77 quatlinrefine, 2d/quatlinrefine.m4:13-114,
// 3d/quatlinrefine.m4).  Both normalise the result (quatnorm, source/fortran/
// quat.f:1043-1083: q / |q|, left as 0 when |q| < 1e-15).  The reference
// passes symmetry-rotation index 0 to quatsymmrotate, which is the identity
// for qlen 1 and 2 (quatsymmrotate2 maps 0 to 1, the identity) and stops the
// run for qlen 4 (quatsymmrotate4 rejects 0, quat.f:325-328); these kernels
// apply the identity for every qlen, the reference's qlen 1 / 2 behaviour.
// Arithmetic in the Fortran order: the coarse sum accumulates q dVf over the
// children (k, then j, then i fastest) and scales by 1/dVc; the refined value
// is the (bi/tri)linear form of quatlinrefine with x = ((r + 1/2) dxf -
// dxc/2) / dxc, shifted down one coarse cell when negative.
#include "mg.h"

namespace ampe {

namespace {

__device__ __forceinline__ void qnorm(double* q, int ql)
{
   double s = 0.0;
   for (int m = 0; m < ql; m++) s = s + q[m] * q[m];
   double mg = sqrt(s);
   double minv = mg < 1.0e-15 ? 0.0 : 1.0 / mg;
   for (int m = 0; m < ql; m++) q[m] = q[m] * minv;
}

constexpr int kQl = 4;

template <int NDIM>
__global__ void __launch_bounds__(256) k_quat_coarsen(LevelArgs ca, LevelArgs fa, int ql,
                                                      double dVf, double dVcinv,
                                                      const double* __restrict__ qf,
                                                      double* __restrict__ qc)
{
   CellPos p;
   if (!cell_pos(ca, 0, p)) return;
   const PatchDev fp = fa.P[ca.T[tile_index(ca)].patch];
   double acc[kQl] = {0.0, 0.0, 0.0, 0.0};
   const int nz = NDIM == 3 ? 2 : 1;
   for (int r2 = 0; r2 < nz; r2++)
      for (int r1 = 0; r1 < 2; r1++)
         for (int r0 = 0; r0 < 2; r0++) {
            int fi = 2 * p.i + r0, fj = 2 * p.j + r1, fk = NDIM == 3 ? 2 * p.k + r2 : 0;
            long long f0 = fp.cell_off + (fi + fa.g) + (long long)(fj + fa.g) * fp.sy +
                           (long long)(fk + fa.gz) * fp.sz;
            for (int m = 0; m < ql; m++) acc[m] = acc[m] + qf[m * fa.ncg + f0] * dVf;
         }
   double q[kQl];
   for (int m = 0; m < ql; m++) q[m] = acc[m] * dVcinv;
   qnorm(q, ql);
   for (int m = 0; m < ql; m++) qc[m * ca.ncg + p.c0] = q[m];
}

// coarse index and weight of a fine index (global), quatlinrefine's ic / x
__device__ __forceinline__ void lin_index(int gf, double dxf, double dxc, int& gc, double& x)
{
   int ic = gf < 0 ? (gf + 1) / 2 - 1 : gf / 2;
   int ir = gf - ic * 2;
   double dlt = ((double)ir + 0.5) * dxf - dxc * 0.5;
   x = dlt / dxc;
   if (x < 0.0) {
      ic = ic - 1;
      x = x + 1.0;
   }
   gc = ic;
}

// NORM: quatlinrefine (interpolate, then q/|q|, ql <= kQl); otherwise the plain
// cell-centred linear refinement of every component (the interpolation
// without the normalisation).
template <int NDIM, bool NORM>
__global__ void __launch_bounds__(256) k_quat_refine(LevelArgs fa, LevelArgs ca, int ql,
                                                     double dxf0, double dxf1, double dxf2,
                                                     double dxc0, double dxc1, double dxc2,
                                                     const double* __restrict__ qc,
                                                     double* __restrict__ qf)
{
   CellPos p;
   if (!cell_pos(fa, 0, p)) return;
   const PatchDev cp = ca.P[fa.T[tile_index(fa)].patch];
   int g0, g1, g2 = 0;
   double x, y, z = 0.0;
   lin_index(p.pd.lo[0] + p.i, dxf0, dxc0, g0, x);
   lin_index(p.pd.lo[1] + p.j, dxf1, dxc1, g1, y);
   if (NDIM == 3) lin_index(p.pd.lo[2] + p.k, dxf2, dxc2, g2, z);
   const int c0i = g0 - cp.lo[0], c1i = g1 - cp.lo[1], c2i = NDIM == 3 ? g2 - cp.lo[2] : 0;
   auto at = [&](int a, int b, int c) -> long long {
      return cp.cell_off + (c0i + a + ca.g) + (long long)(c1i + b + ca.g) * cp.sy +
             (long long)(c2i + c + ca.gz) * cp.sz;
   };
   double q[kQl];
   for (int m = 0; m < ql; m++) {
      const double* a = qc + m * ca.ncg;
      double v;
      double q1 = a[at(0, 0, 0)], q2 = a[at(0, 1, 0)], q3 = a[at(1, 0, 0)],
             q4 = a[at(1, 1, 0)];
      if (NDIM == 2) {
         v = (q1 + (q3 - q1) * x) * (1.0 - y) + (q2 + (q4 - q2) * x) * y;
      } else {
         double q5 = a[at(0, 0, 1)], q6 = a[at(0, 1, 1)], q7 = a[at(1, 0, 1)],
                q8 = a[at(1, 1, 1)];
         v = ((q1 + (q3 - q1) * x) * (1.0 - y) + (q2 + (q4 - q2) * x) * y) * (1.0 - z) +
             ((q5 + (q7 - q5) * x) * (1.0 - y) + (q6 + (q8 - q6) * x) * y) * z;
      }
      if (NORM)
         q[m < kQl ? m : kQl - 1] = v;
      else
         qf[m * fa.ncg + p.c0] = v;
   }
   if (NORM) {
      qnorm(q, ql);
      for (int m = 0; m < ql; m++) qf[m * fa.ncg + p.c0] = q[m];
   }
}

}  // namespace

namespace {

int refine_launch(const Level& fine, const Level& coarse, const double* dxf, int ql, bool norm,
                  const double* qc, double* qf, hipStream_t s)
{
   double c[3] = {2.0 * dxf[0], 2.0 * dxf[1], 2.0 * dxf[2]};
   auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(fine.ncell_tiles), dim3(TX, TY), 0, s, level_args(fine),
                         level_args(coarse), ql, dxf[0], dxf[1], dxf[2], c[0], c[1], c[2], qc,
                         qf);
   };
   if (fine.ndim == 3)
      norm ? go(k_quat_refine<3, true>) : go(k_quat_refine<3, false>);
   else
      norm ? go(k_quat_refine<2, true>) : go(k_quat_refine<2, false>);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

}  // namespace

int quat_coarsen(const Level& coarse, const Level& fine, const double* dxf, int ql,
                 const double* qf, double* qc, hipStream_t s)
{
   if (ql < 1 || ql > kQl) {
      set_error("quat_coarsen: qlen must be 1 to 4");
      return -1;
   }
   double dVf = dxf[0] * dxf[1], dVc = (2.0 * dxf[0]) * (2.0 * dxf[1]);
   if (coarse.ndim == 3) {
      dVf = dVf * dxf[2];
      dVc = dVc * (2.0 * dxf[2]);
   }
   double dVcinv = 1.0 / dVc;
   if (coarse.ndim == 3)
      hipLaunchKernelGGL(k_quat_coarsen<3>, dim3(coarse.ncell_tiles), dim3(TX, TY), 0, s,
                         level_args(coarse), level_args(fine), ql, dVf, dVcinv, qf, qc);
   else
      hipLaunchKernelGGL(k_quat_coarsen<2>, dim3(coarse.ncell_tiles), dim3(TX, TY), 0, s,
                         level_args(coarse), level_args(fine), ql, dVf, dVcinv, qf, qc);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int quat_refine(const Level& fine, const Level& coarse, const double* dxf, int ql,
                const double* qc, double* qf, hipStream_t s)
{
   if (ql < 1 || ql > kQl) {
      set_error("quat_refine: qlen must be 1 to 4");
      return -1;
   }
   return refine_launch(fine, coarse, dxf, ql, true, qc, qf, s);
}

int linear_refine(const Level& fine, const Level& coarse, const double* dxf, int ncomp,
                  const double* c, double* f, hipStream_t s)
{
   if (ncomp < 1) {
      set_error("linear_refine: ncomp must be >= 1");
      return -1;
   }
   return refine_launch(fine, coarse, dxf, ncomp, false, c, f, s);
}

}  // namespace ampe
