// Right-hand-side kernels of the phase, orientation and composition
// equations, the per-cell KKS solve, the q projection and AMR tagging.
//
// Every kernel processes a whole level (all patches) in one launch over the
// level's tile list.  Face quantities (fluxes, face coefficients) are
// recomputed per cell from the stencil instead of being materialised as side
// arrays, so each kernel reads its cell fields once and writes its outputs
// once; the arithmetic of each face is the reference's, in its order.
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ampe {

// Single precision literal 1.e-12 of anisotropic_gradient_flux promoted to
// double (source/fortran/2d/quatrhs.m4:193, 3d/quatrhs.m4:184).
#define AMPE_ANISO_THRESHOLD ((double)1.e-12f)

// Launch-uniform quantities derived from the model on the host: IEEE division
// gives the device's value, and the kernels do no per-cell division or exp
// for them.
RhsConst rhs_const(const ampe_model_params& p)
{
   RhsConst k;
   double e2 = p.epsilon_phi * p.epsilon_phi;
   for (int d = 0; d < 3; d++) {
      k.dxinv[d] = 1.0 / p.dx[d];
      k.hdxinv[d] = 0.5 / p.dx[d];
      k.e2dx[d] = e2 / p.dx[d];
      k.e2dx12[d] = (1.0 / 12.0) * e2 / p.dx[d];
   }
   k.convL = 1.0e-6 / p.vm_L;
   k.convA = 1.0e-6 / p.vm_A;
   k.maxgni = 1.0 / p.gradient_floor;
   // concentration_pfmdiffusion at the uniform temperature (face T = (T+T)/2)
   double invT = 2.0 / (p.temperature + p.temperature);
   k.Dl = p.D_liquid * std::exp(-(p.Q0_liquid / kRGas) * invT);
   k.Ds = p.D_solid * std::exp(-(p.Q0_solid / kRGas) * invT);
   return k;
}

// ------------------------------------------------------------ phase flux
// Face between cells cm = c0 - st[d] and c0, direction d.
// Simple: gradient_flux (source/fortran/2d/quatrhs.m4:12-53).
// Isotropic (2-D): compute_flux_isotropic (2d/quatrhs.m4:106-151).
// Anisotropic: anisotropic_gradient_flux (2d/quatrhs.m4:154-256,
// 3d/quatrhs.m4:149-356).
__device__ __forceinline__ void quatmult4(const double* a, const double* b, double* r)
{
   r[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
   r[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
   r[2] = a[0] * b[2] + a[2] * b[0] + a[3] * b[1] - a[1] * b[3];
   r[3] = a[0] * b[3] + a[3] * b[0] + a[1] * b[2] - a[2] * b[1];
}

// FLUX < 0 reads p.flux_type at run time.
template <int NDIM, int FLUX = -1>
__device__ double phase_face_flux(const ampe_model_params& p, const RhsConst& k,
                                  const double* __restrict__ ph, const double* __restrict__ q,
                                  long long ncg, const long long* st, int d, long long cm,
                                  long long c0)
{
   const int ft = (FLUX < 0) ? p.flux_type : FLUX;
   if (ft == AMPE_FLUX_ANISOTROPIC) {
      if (NDIM == 2) {
         double eps = p.epsilon_phi, nu = p.aniso_nu;
         int kn = p.aniso_knumber;
         double pi = 4.0 * atan(1.0);
         double dinv0 = k.dxinv[0], dinv1 = k.dxinv[1];
         int t = 1 - d;  // transverse direction
         double gn = (ph[c0] - ph[cm]) * (d == 0 ? dinv0 : dinv1);
         double gt = 0.25 * (ph[cm + st[t]] - ph[cm - st[t]] + ph[c0 + st[t]] - ph[c0 - st[t]]) *
                     (d == 0 ? dinv1 : dinv0);
         double dphidx = (d == 0) ? gn : gt;
         double dphidy = (d == 0) ? gt : gn;
         double theta = (fabs(dphidx) > AMPE_ANISO_THRESHOLD) ? atan(dphidy / dphidx) : 0.5 * pi;
         double qa = 0.5 * (q[cm] + q[c0]);
         if (qa > 1.0) qa = 1.0;
         if (qa < -1.0) qa = -1.0;
         double phia = (p.qlen == 4) ? 2.0 * acos(qa) : acos(qa);
         double epst = eps * (1.0 + nu * cos(kn * (theta - phia)));
         double deps = -kn * eps * nu * sin(kn * (theta - phia));
         if (d == 0) return epst * epst * dphidx - epst * deps * dphidy;
         return epst * epst * dphidy + epst * deps * dphidx;
      } else {
         double eps = p.epsilon_phi, eps4 = p.aniso_nu;
         const double* dinv = k.dxinv;
         double factor = 4.0 * eps4 / (1.0 - 3.0 * eps4);
         double gr[3];
#pragma unroll
         for (int e = 0; e < 3; e++) {
            if (e == d)
               gr[e] = (ph[c0] - ph[cm]) * dinv[e];
            else
               gr[e] = 0.25 * (ph[cm + st[e]] - ph[cm - st[e]] + ph[c0 + st[e]] - ph[c0 - st[e]]) *
                       dinv[e];
         }
         double gphi2 = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
         double dg[4], n4;
         if (fabs(gphi2) > AMPE_ANISO_THRESHOLD) {
            double nni = 1.0 / sqrt(gphi2);
            double n[4] = {0.0, gr[0] * nni, gr[1] * nni, gr[2] * nni};
            double qa[4];
#pragma unroll
            for (int m = 0; m < 4; m++) qa[m] = 0.5 * (q[m * ncg + cm] + q[m * ncg + c0]);
            double qp[4] = {qa[0], -qa[1], -qa[2], -qa[3]}, tq[4], np[4];
            quatmult4(n, qp, tq);
            quatmult4(qa, tq, np);
            double a1 = np[1] * np[1], a2 = np[2] * np[2], a3 = np[3] * np[3];
            n4 = a1 * a1 + a2 * a2 + a3 * a3;
            double dd[4] = {0.0, np[1] * (np[1] * np[1] - n4), np[2] * (np[2] * np[2] - n4),
                            np[3] * (np[3] * np[3] - n4)};
            quatmult4(dd, qa, tq);
            quatmult4(qp, tq, dg);
         } else {
            dg[0] = dg[1] = dg[2] = dg[3] = 0.0;
            n4 = 0.0;
         }
         double gamma = eps * (1.0 - 3.0 * eps4) * (1.0 + factor * n4);
         return gamma * gamma * gr[d] + 16.0 * eps * gamma * eps4 * sqrt(gphi2) * dg[d + 1];
      }
   }
   if (ft == AMPE_FLUX_ISOTROPIC && NDIM == 2) {
      double dinv = k.e2dx12[d];
      long long t = st[1 - d];
      return dinv * ((ph[c0 - t] - ph[cm - t]) + (ph[c0] - ph[cm]) * 10.0 + (ph[c0 + t] - ph[cm + t]));
   }
   return (ph[c0] - ph[cm]) * k.e2dx[d];
}

// |grad q| from cell-centred gradients: quatdiffs -> quatgrad_cell ->
// quatgrad_modulus (2d/quatgrad.m4:12-60, 882-915).
template <int NDIM>
__device__ __forceinline__ double quat_cell_modulus(const ampe_model_params& p, const RhsConst& k,
                                                    const double* __restrict__ q, long long ncg,
                                                    const long long* st, long long c0)
{
   const double* p5 = k.hdxinv;
   double s = 0.0;
   for (int m = 0; m < p.qlen; m++) {
      const double* a = q + m * ncg;
      double gx = ((a[c0 + st[0]] - a[c0]) + (a[c0] - a[c0 - st[0]])) * p5[0];
      double gy = ((a[c0 + st[1]] - a[c0]) + (a[c0] - a[c0 - st[1]])) * p5[1];
      if (NDIM == 3) {
         double gz = ((a[c0 + st[2]] - a[c0]) + (a[c0] - a[c0 - st[2]])) * p5[2];
         s = s + gx * gx + gy * gy + gz * gz;
      } else {
         s = s + gx * gx + gy * gy;
      }
   }
   return sqrt(s);
}

// ------------------------------------------------------------ quat grads
template <int NDIM>
__global__ void __launch_bounds__(256) k_quat_modulus(LevelArgs a, ampe_model_params p,
                                                      RhsConst k, const double* __restrict__ q,
                                                      double* __restrict__ gmod)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   gmod[c.c0] = quat_cell_modulus<NDIM>(p, k, q, a.ncg, c.st, c.c0);
}

// |grad q|^2 on faces from the full side gradient (quatgrad_side
// 2d/quatgrad.m4:157-229, summed as compute_face_coef2d quatfacops.m4:83-92).
// QL > 0: the component count as a compile-time constant (loops unrolled);
// QL = 0 reads p.qlen.
template <int NDIM, int QL = 0>
__device__ __forceinline__ double quat_face_norm2(const ampe_model_params& p, const RhsConst& k,
                                                  const double* __restrict__ q, long long ncg,
                                                  const long long* st, int d, long long cm,
                                                  long long c0)
{
   double s = 0.0;
   const int ql = QL > 0 ? QL : p.qlen;
#pragma unroll
   for (int n = 0; n < NDIM; n++) {
      double dinv = k.dxinv[n];
      double p25 = 0.25 * dinv;
#pragma unroll
      for (int m = 0; m < ql; m++) {
         const double* aa = q + m * ncg;
         double gv;
         if (n == d) {
            gv = dinv * (aa[c0] - aa[cm]);
         } else {
            long long t = st[n];
            gv = p25 * ((aa[cm + t] - aa[cm]) + (aa[cm] - aa[cm - t]) + (aa[c0 + t] - aa[c0]) +
                        (aa[c0] - aa[c0 - t]));
         }
         s = s + gv * gv;
      }
   }
   return s;
}

template <int NDIM, int QL>
__global__ void __launch_bounds__(256) k_quat_face_norm2(LevelArgs a, ampe_model_params p,
                                                         RhsConst k, const double* __restrict__ q,
                                                         double* __restrict__ n2)
{
   CellPos c;
   if (!cell_pos(a, 1, c)) return;
   if (NDIM == 2 && c.k > 0) return;
   for (int d = 0; d < NDIM; d++) {
      bool ok = (c.i < c.pd.nx + (d == 0)) && (c.j < c.pd.ny + (d == 1)) &&
                (NDIM == 2 || c.k < c.pd.nz + (d == 2));
      if (!ok) continue;
      n2[face_idx(a, c.pd, d, c.i, c.j, c.k)] =
          quat_face_norm2<NDIM, QL>(p, k, q, a.ncg, c.st, d, c.c0 - c.st[d], c.c0);
   }
}

// The same face norms in 3-D with the workgroup's q values staged in LDS: the
// tile's cells grown by one in each direction (planes k-1, k, k+1) are read
// once, coalesced along i, and every face of the tile takes its values from
// LDS in quat_face_norm2's order (the global-memory form reads up to ten
// neighbours per component per face).  Halo cells outside the ghost box are
// only needed by faces that do not exist; they are set to zero.
template <int QL>
__global__ void __launch_bounds__(256) k_quat_face_norm2_lds(LevelArgs a, ampe_model_params p,
                                                             RhsConst k,
                                                             const double* __restrict__ q,
                                                             double* __restrict__ n2)
{
   constexpr int HX = TX + 2, HY = TY + 2;
   __shared__ double sq[QL][3][HY][HX];
   const Tile t = a.T[tile_index(a)];
   const PatchDev pd = a.P[t.patch];
   const int tid = (int)threadIdx.y * TX + (int)threadIdx.x;
   for (int e = tid; e < 3 * HY * HX; e += TX * TY) {
      const int ii = e % HX, r = e / HX;
      const int jj = r % HY, kk = r / HY;
      const int ci = t.i0 - 1 + ii, cj = t.j0 - 1 + jj, ck = t.k - 1 + kk;
      const bool ok = ci >= -a.g && ci < pd.nx + a.g && cj >= -a.g && cj < pd.ny + a.g &&
                      ck >= -a.gz && ck < pd.nz + a.gz;
      const long long idx = pd.cell_off + (ci + a.g) + (long long)(cj + a.g) * pd.sy +
                            (long long)(ck + a.gz) * pd.sz;
#pragma unroll
      for (int m = 0; m < QL; m++) sq[m][kk][jj][ii] = ok ? q[m * a.ncg + idx] : 0.0;
   }
   __syncthreads();
   const int i = t.i0 + (int)threadIdx.x, j = t.j0 + (int)threadIdx.y, kc = t.k;
   if (i >= pd.nx + 1 || j >= pd.ny + 1) return;
   const int x0 = (int)threadIdx.x + 1, y0 = (int)threadIdx.y + 1;
#pragma unroll
   for (int d = 0; d < 3; d++) {
      const bool ok = (i < pd.nx + (d == 0)) && (j < pd.ny + (d == 1)) && (kc < pd.nz + (d == 2));
      if (!ok) continue;
      // cm = the cell across the low face of direction d
      const int mx = x0 - (d == 0), my = y0 - (d == 1), mz = 1 - (d == 2);
      double s = 0.0;
#pragma unroll
      for (int n = 0; n < 3; n++) {
         const double dinv = k.dxinv[n];
         const double p25 = 0.25 * dinv;
         const int tx = (n == 0), ty = (n == 1), tz = (n == 2);
#pragma unroll
         for (int m = 0; m < QL; m++) {
            const double c0v = sq[m][1][y0][x0], cmv = sq[m][mz][my][mx];
            double gv;
            if (n == d) {
               gv = dinv * (c0v - cmv);
            } else {
               gv = p25 * ((sq[m][mz + tz][my + ty][mx + tx] - cmv) +
                           (cmv - sq[m][mz - tz][my - ty][mx - tx]) +
                           (sq[m][1 + tz][y0 + ty][x0 + tx] - c0v) +
                           (c0v - sq[m][1 - tz][y0 - ty][x0 - tx]));
            }
            s = s + gv * gv;
         }
      }
      n2[face_idx(a, pd, d, i, j, kc)] = s;
   }
}

// ------------------------------------------------------------ phase RHS
// PhaseRHSStrategyWithQ::evaluateRHS per patch (source/PhaseRHSStrategyWithQ.cc
// :151-312): flux divergence, double well, orientation terms (computerhspbg,
// 2d/quatrhs.m4:264-405), KKS driving force (QuadraticFreeEnergyStrategy.cc
// :393-511 / CALPHADFreeEnergyStrategyBinary.cc:521-651), times M_phi.
// Specialised on the flux type FLUX and the free-energy model FE (kFeNone: no
// composition field) so that each instance holds only its own path's registers.
constexpr int kFeNone = 3;

template <int NDIM, int FE>
__device__ __forceinline__ void phase_rhs_tail(const LevelArgs& a, const ampe_model_params& p,
                                               const RhsConst& k, const CellPos& c, double r,
                                               const double* __restrict__ phi,
                                               const double* __restrict__ q,
                                               const double* __restrict__ cl,
                                               const double* __restrict__ cs,
                                               double* __restrict__ rhs);

template <int NDIM, int FLUX, int FE>
__device__ __forceinline__ void phase_rhs_cell(const LevelArgs& a, const ampe_model_params& p,
                                               const RhsConst& k, const CellPos& c,
                                               const double* __restrict__ phi,
                                               const double* __restrict__ q,
                                               const double* __restrict__ cl,
                                               const double* __restrict__ cs,
                                               double* __restrict__ rhs)
{
   const long long c0 = c.c0;
   double r = 0.0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      double flo = phase_face_flux<NDIM, FLUX>(p, k, phi, q, a.ncg, c.st, d, c0 - c.st[d], c0);
      double fhi = phase_face_flux<NDIM, FLUX>(p, k, phi, q, a.ncg, c.st, d, c0, c0 + c.st[d]);
      double dt = (fhi - flo) * k.dxinv[d];
      r = (d == 0) ? dt : r + dt;
   }
   phase_rhs_tail<NDIM, FE>(a, p, k, c, r, phi, q, cl, cs, rhs);
}

// The phase RHS of one cell after the flux divergence r: double well,
// orientation terms, driving force, times the mobility.
template <int NDIM, int FE>
__device__ __forceinline__ void phase_rhs_tail(const LevelArgs& a, const ampe_model_params& p,
                                               const RhsConst& k, const CellPos& c, double r,
                                               const double* __restrict__ phi,
                                               const double* __restrict__ q,
                                               const double* __restrict__ cl,
                                               const double* __restrict__ cs,
                                               double* __restrict__ rhs)
{
   const long long c0 = c.c0;
   double ph = phi[c0];
   r = r - p.phi_well_scale * dwell(ph, p.phi_well_type);
   if (p.evolve_quat) {
      double mf = 2.0 * p.H_parameter, T = p.temperature;
      double eq2 = 0.5 * p.epsilon_q * p.epsilon_q;
      double p1 = dinterp(ph, p.orient_interp1);
      double p2 = dinterp(ph, p.orient_interp2);
      double gm = quat_cell_modulus<NDIM>(p, k, q, a.ncg, c.st, c0);
      r = r - mf * T * p1 * gm - p2 * eq2 * gm * gm;
   }
   double mob = p.phi_mobility;
   if (FE != kFeNone) {
      double convL = k.convL, convA = k.convA;
      double c_l = cl[c0], c_a = cs[c0];
      double f_l = fenergy_t<FE>(p, c_l, 0) * convL;
      double f_a = fenergy_t<FE>(p, c_a, 1) * convA;
      double mu = (FE == AMPE_FE_QUADRATIC) ? dfenergy_t<FE>(p, c_l, 0) * convL
                                            : dfenergy_t<FE>(p, c_a, 1) * convA;
      double hp = dinterp(ph, p.energy_interp);
      r += hp * ((f_l - f_a) - mu * (c_l - c_a));
      if (p.phi_mobility_type == AMPE_PHI_MOB_KIM) mob = phase_mobility_t<FE>(p, c_l, c_a);
   }
   rhs[c0] = mob * r;
}

// Minimum waves per SIMD the register allocation of k_phase_rhs must allow.
#ifndef AMPE_PHASE_RHS_WAVES
#define AMPE_PHASE_RHS_WAVES 1
#endif

template <int NDIM, int FLUX, int FE>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(AMPE_PHASE_RHS_WAVES)))
    k_phase_rhs(LevelArgs a, ampe_model_params p, RhsConst k,
                                                   const double* __restrict__ phi,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ cl,
                                                   const double* __restrict__ cs,
                                                   double* __restrict__ rhs)
{
   for (int kk = 0; kk < a.nz_tile; kk++) {
      CellPos c;
      if (!cell_pos_col(a, kk, c)) {
         if (c.i >= c.pd.nx || c.j >= c.pd.ny) return;
         break;
      }
      phase_rhs_cell<NDIM, FLUX, FE>(a, p, k, c, phi, q, cl, cs, rhs);
   }
}

// The phase RHS with each x and y face flux evaluated once: a workgroup's
// tile (one plane) puts the fluxes of its cells' low faces, and of the high
// face closing each row and column, in LDS, then every cell takes the
// divergence from there (z faces, 3-D: both per cell).  The per-cell form
// evaluates every face twice, once from each side; the anisotropic flux has
// about four transcendentals per face.  Same values as phase_rhs_cell.
template <int NDIM, int FLUX, int FE>
__global__ void __launch_bounds__(256) k_phase_rhs_faces(LevelArgs a, ampe_model_params p,
                                                         RhsConst k,
                                                         const double* __restrict__ phi,
                                                         const double* __restrict__ q,
                                                         const double* __restrict__ cl,
                                                         const double* __restrict__ cs,
                                                         double* __restrict__ rhs)
{
   __shared__ double fx[TY][TX + 1];
   __shared__ double fy[TY + 1][TX];
   const Tile t = a.T[tile_index(a)];
   CellPos c;
   c.pd = a.P[t.patch];
   const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
   const int nloc = min(TX, c.pd.nx - t.i0), mloc = min(TY, c.pd.ny - t.j0);
   c.i = t.i0 + tx;
   c.j = t.j0 + ty;
   c.k = t.k;
   c.st[0] = 1;
   c.st[1] = c.pd.sy;
   c.st[2] = c.pd.sz;
   c.c0 = c.pd.cell_off + (c.i + a.g) + (long long)(c.j + a.g) * c.pd.sy +
          (long long)(c.k + a.gz) * c.pd.sz;
   const long long c0 = c.c0, sy = c.st[1];
   if (ty < mloc) {
      if (tx < nloc) fx[ty][tx] = phase_face_flux<NDIM, FLUX>(p, k, phi, q, a.ncg, c.st, 0, c0 - 1, c0);
      if (tx == 0) {
         const long long e = c0 + nloc;
         fx[ty][nloc] = phase_face_flux<NDIM, FLUX>(p, k, phi, q, a.ncg, c.st, 0, e - 1, e);
      }
   }
   if (tx < nloc) {
      if (ty < mloc) fy[ty][tx] = phase_face_flux<NDIM, FLUX>(p, k, phi, q, a.ncg, c.st, 1, c0 - sy, c0);
      if (ty == 0) {
         const long long e = c0 + (long long)mloc * sy;
         fy[mloc][tx] = phase_face_flux<NDIM, FLUX>(p, k, phi, q, a.ncg, c.st, 1, e - sy, e);
      }
   }
   __syncthreads();
   if (tx >= nloc || ty >= mloc) return;
   double r = (fx[ty][tx + 1] - fx[ty][tx]) * k.dxinv[0];
   r = r + (fy[ty + 1][tx] - fy[ty][tx]) * k.dxinv[1];
   if (NDIM == 3) {
      const long long sz = c.st[2];
      double flo = phase_face_flux<NDIM, FLUX>(p, k, phi, q, a.ncg, c.st, 2, c0 - sz, c0);
      double fhi = phase_face_flux<NDIM, FLUX>(p, k, phi, q, a.ncg, c.st, 2, c0, c0 + sz);
      r = r + (fhi - flo) * k.dxinv[2];
   }
   phase_rhs_tail<NDIM, FE>(a, p, k, c, r, phi, q, cl, cs, rhs);
}

// ------------------------------------------------------------ quat RHS
// compute_face_coef2d/3d (2d/quatfacops.m4:14-122) from the face |grad q|^2
__device__ __forceinline__ double quat_face_coef(const ampe_model_params& p, const RhsConst& k,
                                                 double phm, double ph0, double n2f)
{
   double floor2 = p.gradient_floor * p.gradient_floor;
   double maxgni = k.maxgni;
   double eps2 = p.epsilon_q * p.epsilon_q;
   double mf = 2.0 * p.H_parameter;
   double phia = face_avg(phm, ph0, p.avg_type);
   double tempa = 0.5 * (p.temperature + p.temperature);
   double diff = mf * tempa * interp(phia, p.orient_interp1);
   double hphi2 = interp(phia, p.orient_interp2);
   double gni = grad_normi(n2f, p.floor_type, floor2, maxgni);
   return -gni * diff - eps2 * hphi2;
}

// QuatFACOps::evaluateRHS with use_gradq_for_flux and projection
// (source/QuatFACOps.cc:1861-1890, 2089-2108): flux = fc dq/dh
// (compute_flux2d_from_gradq), lambda (compute_lambda_flux2d/3d),
// rhs = -M_q (div + 2 q lambda) (add_quat_proj_op2d/3d).
// FC: `face` holds the face coefficients (one evaluation per face by
// k_quat_face_coef); otherwise it holds the face |grad q|^2 and each cell
// evaluates the coefficients of its six faces.
template <int NDIM, bool FC>
__global__ void __launch_bounds__(256) k_quat_rhs(LevelArgs a, ampe_model_params p, RhsConst k,
                                                  const double* __restrict__ phi,
                                                  const double* __restrict__ q,
                                                  const double* __restrict__ face,
                                                  double* __restrict__ rhs)
{
   AMPE_COLUMN_BEGIN(a, c)
   const long long c0 = c.c0, ncg = a.ncg;
   const int ql = p.qlen;
   double fclo[3], fchi[3];
   const int ijk[3] = {c.i, c.j, c.k};
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      int lo[3] = {ijk[0], ijk[1], ijk[2]};
      long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      lo[d] += 1;
      long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      if (FC) {
         fclo[d] = face[flo];
         fchi[d] = face[fhi];
      } else {
         fclo[d] = quat_face_coef(p, k, phi[c0 - c.st[d]], phi[c0], face[flo]);
         fchi[d] = quat_face_coef(p, k, phi[c0], phi[c0 + c.st[d]], face[fhi]);
      }
   }
   const double* dinv = k.dxinv;
   const double* fac = k.hdxinv;
   double lam = 0.0, sumq2 = 0.0;
   double dv[4];
   for (int m = 0; m < ql; m++) {
      const double* aa = q + m * ncg;
      double qm = aa[c0];
      double dif[3];
#pragma unroll
      for (int d = 0; d < NDIM; d++) {
         long long s = c.st[d];
         double left = fclo[d] * (dinv[d] * (aa[c0] - aa[c0 - s]));
         double right = fchi[d] * (dinv[d] * (aa[c0 + s] - aa[c0]));
         dif[d] = right - left;
      }
      if (NDIM == 3) {
         lam = lam - qm * (dif[0] * fac[0] + dif[1] * fac[1] + dif[2] * fac[2]);
         dv[m] = dif[0] * dinv[0] + dif[1] * dinv[1] + dif[2] * dinv[2];
      } else {
         lam = lam - qm * (dif[0] * fac[0] + dif[1] * fac[1]);
         dv[m] = dif[0] * dinv[0] + dif[1] * dinv[1];
      }
      sumq2 = sumq2 + qm * qm;
   }
   lam = lam / sumq2;
   double mq = quat_mobility(p, phi[c0]);
   for (int m = 0; m < ql; m++) {
      double qm = q[m * ncg + c0];
      rhs[m * ncg + c0] = 0.0 - mq * (dv[m] + 2.0 * qm * lam);
   }
AMPE_COLUMN_END
}

// project2d/3d (2d/quatfacops.m4:940-997)
__global__ void __launch_bounds__(256) k_quat_project(LevelArgs a, int ql,
                                                      const double* __restrict__ q,
                                                      double* __restrict__ corr,
                                                      double* __restrict__ err)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   const long long c0 = c.c0, ncg = a.ncg;
   double qv[4], cv[4];
   double fac = 0.0;
   for (int m = 0; m < ql; m++) {
      qv[m] = q[m * ncg + c0];
      fac = fac + qv[m] * qv[m];
   }
   fac = 1.0 / sqrt(fac);
   for (int m = 0; m < ql; m++) cv[m] = qv[m] * fac;
   fac = 0.0;
   for (int m = 0; m < ql; m++) fac = fac + cv[m] * err[m * ncg + c0];
   for (int m = 0; m < ql; m++) err[m * ncg + c0] = err[m * ncg + c0] - cv[m] * fac;
   for (int m = 0; m < ql; m++) corr[m * ncg + c0] = cv[m] - qv[m];
}

// ------------------------------------------------------------ composition
// concentration_pfmdiffusion (2d/concentrationdiffusion.m4:14-96) and
// setDiffCoeffForPhaseOnPatch (source/KKSCompositionRHSStrategy.cc:210-373)
template <int NDIM>
__global__ void __launch_bounds__(256) k_conc_coeffs(LevelArgs a, ampe_model_params p, RhsConst k,
                                                     const double* __restrict__ phi,
                                                     const double* __restrict__ cl,
                                                     const double* __restrict__ cs,
                                                     double* __restrict__ Dc,
                                                     double* __restrict__ Dphi)
{
   CellPos c;
   if (!cell_pos(a, 1, c)) return;
   if (NDIM == 2 && c.k > 0) return;
   for (int d = 0; d < NDIM; d++) {
      bool ok = (c.i < c.pd.nx + (d == 0)) && (c.j < c.pd.ny + (d == 1)) &&
                (NDIM == 2 || c.k < c.pd.nz + (d == 2));
      if (!ok) continue;
      long long c0 = c.c0, cm = c0 - c.st[d];
      long long fi = face_idx(a, c.pd, d, c.i, c.j, c.k);
      double vphi = face_avg(phi[cm], phi[c0], p.avg_type);
      double hphi = interp(vphi, p.energy_interp);
      double dl = k.Dl, ds = k.Ds;
      double dd = (1.0 - hphi) * dl + hphi * ds;
      Dc[fi] = dd;
      double ph = face_avg(phi[c0], phi[cm], p.avg_type);
      double c_l = 0.5 * (cl[c0] + cl[cm]);
      double c_a = 0.5 * (cs[c0] + cs[cm]);
      double hp = dinterp(ph, p.energy_interp);
      Dphi[fi] = dd * hp * (c_l - c_a);
   }
}

// concentrationflux (2d/concentrationrhs.m4:15-82) + computerhsconcentration
// (2d/concentrationrhs.m4:326-367)
template <int NDIM>
__global__ void __launch_bounds__(256) k_conc_rhs(LevelArgs a, ampe_model_params p, RhsConst k,
                                                  const double* __restrict__ cc,
                                                  const double* __restrict__ phi,
                                                  const double* __restrict__ Dc,
                                                  const double* __restrict__ Dphi,
                                                  double* __restrict__ rhs)
{
   AMPE_COLUMN_BEGIN(a, c)
   const long long c0 = c.c0;
   const int ijk[3] = {c.i, c.j, c.k};
   double s = 0.0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      double dinv = k.dxinv[d];
      int lo[3] = {ijk[0], ijk[1], ijk[2]};
      long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      lo[d] += 1;
      long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      long long cm = c0 - c.st[d], cp = c0 + c.st[d];
      double jl = dinv * (Dc[flo] * (cc[c0] - cc[cm]) + Dphi[flo] * (phi[c0] - phi[cm]));
      double jh = dinv * (Dc[fhi] * (cc[cp] - cc[c0]) + Dphi[fhi] * (phi[cp] - phi[c0]));
      double t = dinv * (jh - jl);
      s = (d == 0) ? t : s + t;
   }
   rhs[c0] = p.conc_mobility * s;
AMPE_COLUMN_END
}

// EBS face diffusivities of the two phases:
// concentration_pfmdiffusion_of_temperature (2d/concentrationdiffusion.m4:341-428)
// as TbasedCompositionDiffusionStrategy::setDiffusion calls it
// (source/TbasedCompositionDiffusionStrategy.cc:305-322), uniform temperature
template <int NDIM>
__global__ void __launch_bounds__(256) k_ebs_coeffs(LevelArgs a, ampe_model_params p, RhsConst k,
                                                    const double* __restrict__ phi,
                                                    double* __restrict__ DL,
                                                    double* __restrict__ DA)
{
   CellPos c;
   if (!cell_pos(a, 1, c)) return;
   if (NDIM == 2 && c.k > 0) return;
   for (int d = 0; d < NDIM; d++) {
      bool ok = (c.i < c.pd.nx + (d == 0)) && (c.j < c.pd.ny + (d == 1)) &&
                (NDIM == 2 || c.k < c.pd.nz + (d == 2));
      if (!ok) continue;
      long long c0 = c.c0, cm = c0 - c.st[d];
      long long fi = face_idx(a, c.pd, d, c.i, c.j, c.k);
      double vphi = face_avg(phi[cm], phi[c0], p.avg_type);
      double hphi = interp(vphi, p.diff_interp);
      double dl = k.Dl, ds = k.Ds;
      DL[fi] = (1.0 - hphi) * dl;
      DA[fi] = hphi * ds;
   }
}

// MobilityCompositionDiffusionStrategy::setDiffusion
// (source/MobilityCompositionDiffusionStrategy.cc:97-150): per face, the
// phase concentrations averaged arithmetically (setDiffCoeffInEachPhaseOnPatch,
// :194-329), each phase's M(c) d2f/dc2(c) (computeLocalDiffusionMatrixL/A,
// :57-95, one species: the 1 x 1 product), then the phase weights
// (setPFMDiffOnPatch, :331-611); uniform temperature.
template <int NDIM>
__global__ void __launch_bounds__(256) k_ebs_coeffs_mob(LevelArgs a, ampe_model_params p,
                                                        MobConst mc,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ cl,
                                                        const double* __restrict__ cs,
                                                        double* __restrict__ DL,
                                                        double* __restrict__ DA)
{
   CellPos c;
   if (!cell_pos(a, 1, c)) return;
   if (NDIM == 2 && c.k > 0) return;
   for (int d = 0; d < NDIM; d++) {
      bool ok = (c.i < c.pd.nx + (d == 0)) && (c.j < c.pd.ny + (d == 1)) &&
                (NDIM == 2 || c.k < c.pd.nz + (d == 2));
      if (!ok) continue;
      long long c0 = c.c0, cm = c0 - c.st[d];
      long long fi = face_idx(a, c.pd, d, c.i, c.j, c.k);
      const double cL = 0.5 * (cl[c0] + cl[cm]);
      const double cA = 0.5 * (cs[c0] + cs[cm]);
      const double dl = phase_diffusion_mobility(mc, 0, cL) * d2fenergy_t<-1>(p, cL, 0);
      const double da = phase_diffusion_mobility(mc, 1, cA) * d2fenergy_t<-1>(p, cA, 1);
      double vphi = face_avg(phi[cm], phi[c0], p.avg_type);
      double hphi = interp(vphi, p.diff_interp);
      DL[fi] = (1.0 - hphi) * dl;
      DA[fi] = hphi * da;
   }
}

// EBSCompositionRHSStrategy::computeFluxOnPatch (source/EBSCompositionRHSStrategy.cc
// :183-328): add_flux (2d/flux.m4:11-66) of the liquid then the solid phase
// onto a zero flux, then computerhsconcentration (2d/concentrationrhs.m4:326-367)
template <int NDIM>
__global__ void __launch_bounds__(256) k_conc_rhs_ebs(LevelArgs a, ampe_model_params p, RhsConst k,
                                                      const double* __restrict__ cl,
                                                      const double* __restrict__ cs,
                                                      const double* __restrict__ DL,
                                                      const double* __restrict__ DA,
                                                      double* __restrict__ rhs)
{
   AMPE_COLUMN_BEGIN(a, c)
   const long long c0 = c.c0;
   const int ijk[3] = {c.i, c.j, c.k};
   double s = 0.0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      double dinv = k.dxinv[d];
      int lo[3] = {ijk[0], ijk[1], ijk[2]};
      long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      lo[d] += 1;
      long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      long long cm = c0 - c.st[d], cp = c0 + c.st[d];
      double jl = 0.0;
      jl = jl + dinv * DL[flo] * (cl[c0] - cl[cm]);
      jl = jl + dinv * DA[flo] * (cs[c0] - cs[cm]);
      double jh = 0.0;
      jh = jh + dinv * DL[fhi] * (cl[cp] - cl[c0]);
      jh = jh + dinv * DA[fhi] * (cs[cp] - cs[c0]);
      double t = dinv * (jh - jl);
      s = (d == 0) ? t : s + t;
   }
   rhs[c0] = p.conc_mobility * s;
AMPE_COLUMN_END
}

// The same with add_flux_4th (2d/flux.m4:139-242, 3d/flux.m4:235-420): each
// face flux adds, per phase, D (15/12 (c_h - c_m) - 1/12 (c_hh - c_mm)) / dx over
// the two cells on either side of the face (the divergence is then the
// fourth-order Laplacian (-1, 16, -30, 16, -1) / 12 dx^2); a face on a
// physical boundary (the patch's first or last face on a side whose physb bit
// is set) keeps the second-order D (c_h - c_m) / dx.  Needs two ghost cells.
__device__ __forceinline__ double ebs_4th_flux(double dxi, double w15, double w1, long long h,
                                               long long st, bool bnd, double dl, double da,
                                               const double* __restrict__ cl,
                                               const double* __restrict__ cs)
{
   const long long m = h - st;
   double f = 0.0;
   if (bnd) {
      f = f + dl * dxi * (cl[h] - cl[m]);
      f = f + da * dxi * (cs[h] - cs[m]);
   } else {
      f = f + dl * (w15 * (cl[h] - cl[m]) - w1 * (cl[h + st] - cl[m - st]));
      f = f + da * (w15 * (cs[h] - cs[m]) - w1 * (cs[h + st] - cs[m - st]));
   }
   return f;
}

template <int NDIM>
__global__ void __launch_bounds__(256) k_conc_rhs_ebs_4th(LevelArgs a, ampe_model_params p,
                                                          RhsConst k,
                                                          const double* __restrict__ cl,
                                                          const double* __restrict__ cs,
                                                          const double* __restrict__ DL,
                                                          const double* __restrict__ DA,
                                                          double* __restrict__ rhs)
{
   AMPE_COLUMN_BEGIN(a, c)
   const long long c0 = c.c0;
   const int ijk[3] = {c.i, c.j, c.k};
   const int n[3] = {c.pd.nx, c.pd.ny, c.pd.nz};
   double s = 0.0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      int lo[3] = {ijk[0], ijk[1], ijk[2]};
      long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      lo[d] += 1;
      long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      const bool blo = ijk[d] == 0 && (c.pd.physb >> (2 * d) & 1);
      const bool bhi = ijk[d] == n[d] - 1 && (c.pd.physb >> (2 * d + 1) & 1);
      const double w15 = 15.0 / (12.0 * p.dx[d]), w1 = 1.0 / (12.0 * p.dx[d]);
      double jl = ebs_4th_flux(k.dxinv[d], w15, w1, c0, c.st[d], blo, DL[flo], DA[flo], cl, cs);
      double jh = ebs_4th_flux(k.dxinv[d], w15, w1, c0 + c.st[d], c.st[d], bhi, DL[fhi],
                               DA[fhi], cl, cs);
      double t = k.dxinv[d] * (jh - jl);
      s = (d == 0) ? t : s + t;
   }
   rhs[c0] = p.conc_mobility * s;
   AMPE_COLUMN_END
}

// The same with add_flux_iso (2d/flux.m4:69-135, 3d/flux.m4:103-232): each face
// flux adds, per phase, its transverse neighbour faces (weight 1/12) and the
// face itself (10/12 in 2-D, 8/12 in 3-D) in the reference's order -- 2-D
// +t, 0, -t as (w D) dc; 3-D w (D dc) over x: -y, -z, +y, +z; y: -x, +x, -z, +z;
// z: -x, +x, -y, +y; then the face -- with every face diffusivity evaluated
// from phid (the arithmetic of k_ebs_coeffs).
template <int NDIM>
__device__ __forceinline__ double ebs_iso_flux(const ampe_model_params& p, const RhsConst& k,
                                               const CellPos& c, int d, long long h0,
                                               double w1, double wc,
                                               const double* __restrict__ phid,
                                               const double* __restrict__ cl,
                                               const double* __restrict__ cs)
{
   long long to[4];
   if (NDIM == 2) {
      long long t = c.st[1 - d];
      to[0] = t;
      to[1] = 0;
      to[2] = -t;
   } else {
      const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
      if (d == 0) {
         to[0] = -c.st[a];
         to[1] = -c.st[b];
         to[2] = c.st[a];
         to[3] = c.st[b];
      } else {
         to[0] = -c.st[a];
         to[1] = c.st[a];
         to[2] = -c.st[b];
         to[3] = c.st[b];
      }
   }
   double f = 0.0;
#pragma unroll
   for (int ph = 0; ph < 2; ph++) {
      const double* cc = ph == 0 ? cl : cs;
      if (NDIM == 2) {
#pragma unroll
         for (int t = 0; t < 3; t++) {
            long long h = h0 + to[t], m = h - c.st[d];
            double hphi = interp(face_avg(phid[m], phid[h], p.avg_type), p.diff_interp);
            double dd = ph == 0 ? (1.0 - hphi) * k.Dl : hphi * k.Ds;
            double w = t == 1 ? wc : w1;
            f = f + w * dd * (cc[h] - cc[m]);
         }
      } else {
#pragma unroll
         for (int t = 0; t < 5; t++) {
            long long h = h0 + (t < 4 ? to[t] : 0), m = h - c.st[d];
            double hphi = interp(face_avg(phid[m], phid[h], p.avg_type), p.diff_interp);
            double dd = ph == 0 ? (1.0 - hphi) * k.Dl : hphi * k.Ds;
            double w = t < 4 ? w1 : wc;
            f = f + w * (dd * (cc[h] - cc[m]));
         }
      }
   }
   return f;
}

template <int NDIM>
__global__ void __launch_bounds__(256) k_conc_rhs_ebs_iso(LevelArgs a, ampe_model_params p,
                                                          RhsConst k,
                                                          const double* __restrict__ phid,
                                                          const double* __restrict__ cl,
                                                          const double* __restrict__ cs,
                                                          double* __restrict__ rhs)
{
   AMPE_COLUMN_BEGIN(a, c)
   const long long c0 = c.c0;
   double s = 0.0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      double dinv = k.dxinv[d];
      double w1 = 1.0 / (12.0 * p.dx[d]);
      double wc = (NDIM == 2 ? 10.0 : 8.0) / (12.0 * p.dx[d]);
      double jl = ebs_iso_flux<NDIM>(p, k, c, d, c0, w1, wc, phid, cl, cs);
      double jh = ebs_iso_flux<NDIM>(p, k, c, d, c0 + c.st[d], w1, wc, phid, cl, cs);
      double t = dinv * (jh - jl);
      s = (d == 0) ? t : s + t;
   }
   rhs[c0] = p.conc_mobility * s;
   AMPE_COLUMN_END
}

// QuatMobilityStrategy::computePhaseMobility over interior cells
// (QuatIntegrator::computeMobilities, source/QuatIntegrator.cc:2959-2992)
__global__ void __launch_bounds__(256) k_phase_mobility(LevelArgs a, ampe_model_params p,
                                                        const double* __restrict__ cl,
                                                        const double* __restrict__ cs,
                                                        double* __restrict__ M)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   M[c.c0] = phase_mobility(p, cl ? cl[c.c0] : 0.0, cs ? cs[c.c0] : 0.0);
}

int launch_phase_mobility(const Level& L, const ampe_model_params& p, const double* cl,
                          const double* cs, double* M, hipStream_t s)
{
   hipLaunchKernelGGL(k_phase_mobility, dim3(L.ncell_tiles), dim3(TX, TY), 0, s, level_args(L),
                      p, cl, cs, M);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

// ------------------------------------------------------------ tagging
// tagfromgrads (2d/quattag.m4:9-55) on the cell gradient of phi
template <int NDIM>
__global__ void __launch_bounds__(256) k_tag(LevelArgs a, ampe_model_params p, RhsConst k,
                                             const double* __restrict__ phi, double tol2,
                                             int* __restrict__ tags)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   const long long c0 = c.c0;
   double g = 0.0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      double p5 = k.hdxinv[d];
      long long s = c.st[d];
      double gd = ((phi[c0 + s] - phi[c0]) + (phi[c0] - phi[c0 - s])) * p5;
      double gh = gd * p.dx[d];
      g = g + gh * gh;
   }
   tags[int_idx(c)] = (g >= tol2) ? 1 : 0;
}

// ------------------------------------------------------------ launchers
static dim3 tile_block() { return dim3(TX, TY); }

#define AMPE_DISPATCH(L, KERNEL, GRID, ...)                                    \
   do {                                                                        \
      if ((L).ndim == 3)                                                       \
         hipLaunchKernelGGL(KERNEL<3>, dim3(GRID), tile_block(), 0, s, __VA_ARGS__); \
      else                                                                     \
         hipLaunchKernelGGL(KERNEL<2>, dim3(GRID), tile_block(), 0, s, __VA_ARGS__); \
      AMPE_HIP_CHECK(hipGetLastError());                                       \
   } while (0)

int launch_quat_grad(const Level& L, const ampe_model_params& p, const double* q,
                     double* gmod, double* face_norm2, hipStream_t s)
{
   if (gmod) AMPE_DISPATCH(L, k_quat_modulus, L.ncell_tiles, level_args(L), p, rhs_const(p), q, gmod);
   if (face_norm2) {
      dim3 g(L.nface_tiles), b = tile_block();
      LevelArgs a = level_args(L, true);
      RhsConst k = rhs_const(p);
      if (L.ndim == 3 && p.qlen == 4)
         hipLaunchKernelGGL((k_quat_face_norm2_lds<4>), g, b, 0, s, a, p, k, q, face_norm2);
      else if (L.ndim == 3)
         hipLaunchKernelGGL((k_quat_face_norm2<3, 0>), g, b, 0, s, a, p, k, q, face_norm2);
      else if (p.qlen == 4)
         hipLaunchKernelGGL((k_quat_face_norm2<2, 4>), g, b, 0, s, a, p, k, q, face_norm2);
      else
         hipLaunchKernelGGL((k_quat_face_norm2<2, 0>), g, b, 0, s, a, p, k, q, face_norm2);
      AMPE_HIP_CHECK(hipGetLastError());
   }
   return 0;
}

template <int NDIM, int FLUX>
static void launch_phase_rhs_fe(const Level& L, const ampe_model_params& p, int fe,
                                const double* phi, const double* q, const double* cl,
                                const double* cs, double* rhs, hipStream_t s)
{
   RhsConst k = rhs_const(p);
   if constexpr (FLUX == AMPE_FLUX_ANISOTROPIC) {
      // each face flux once, through LDS
      dim3 g(L.ncell_tiles), b = tile_block();
      LevelArgs a = level_args(L);
      switch (fe) {
      case AMPE_FE_QUADRATIC:
         hipLaunchKernelGGL((k_phase_rhs_faces<NDIM, FLUX, AMPE_FE_QUADRATIC>), g, b, 0, s, a, p,
                            k, phi, q, cl, cs, rhs);
         break;
      case AMPE_FE_CALPHAD_BINARY:
         hipLaunchKernelGGL((k_phase_rhs_faces<NDIM, FLUX, AMPE_FE_CALPHAD_BINARY>), g, b, 0, s, a,
                            p, k, phi, q, cl, cs, rhs);
         break;
      case AMPE_FE_KKS_DILUTE:
         hipLaunchKernelGGL((k_phase_rhs_faces<NDIM, FLUX, AMPE_FE_KKS_DILUTE>), g, b, 0, s, a, p,
                            k, phi, q, cl, cs, rhs);
         break;
      default:
         hipLaunchKernelGGL((k_phase_rhs_faces<NDIM, FLUX, kFeNone>), g, b, 0, s, a, p, k, phi, q,
                            cl, cs, rhs);
      }
   } else {
      dim3 g(L.ncol_tiles), b = tile_block();
      LevelArgs a = level_args_col(L);
      switch (fe) {
      case AMPE_FE_QUADRATIC:
         hipLaunchKernelGGL((k_phase_rhs<NDIM, FLUX, AMPE_FE_QUADRATIC>), g, b, 0, s, a, p, k, phi, q,
                            cl, cs, rhs);
         break;
      case AMPE_FE_CALPHAD_BINARY:
         hipLaunchKernelGGL((k_phase_rhs<NDIM, FLUX, AMPE_FE_CALPHAD_BINARY>), g, b, 0, s, a, p, k, phi,
                            q, cl, cs, rhs);
         break;
      case AMPE_FE_KKS_DILUTE:
         hipLaunchKernelGGL((k_phase_rhs<NDIM, FLUX, AMPE_FE_KKS_DILUTE>), g, b, 0, s, a, p, k, phi, q,
                            cl, cs, rhs);
         break;
      default:
         hipLaunchKernelGGL((k_phase_rhs<NDIM, FLUX, kFeNone>), g, b, 0, s, a, p, k, phi, q, cl, cs,
                            rhs);
      }
   }
}

int launch_phase_rhs(const Level& L, const ampe_model_params& p, const double* phi,
                     const double* q, const double* cl, const double* cs, double* rhs,
                     hipStream_t s)
{
   int fe = p.with_conc ? p.fe_model : kFeNone;
   if (L.ndim == 3) {
      if (p.flux_type == AMPE_FLUX_ANISOTROPIC)
         launch_phase_rhs_fe<3, AMPE_FLUX_ANISOTROPIC>(L, p, fe, phi, q, cl, cs, rhs, s);
      else
         launch_phase_rhs_fe<3, AMPE_FLUX_SIMPLE>(L, p, fe, phi, q, cl, cs, rhs, s);
   } else {
      if (p.flux_type == AMPE_FLUX_ANISOTROPIC)
         launch_phase_rhs_fe<2, AMPE_FLUX_ANISOTROPIC>(L, p, fe, phi, q, cl, cs, rhs, s);
      else if (p.flux_type == AMPE_FLUX_ISOTROPIC)
         launch_phase_rhs_fe<2, AMPE_FLUX_ISOTROPIC>(L, p, fe, phi, q, cl, cs, rhs, s);
      else
         launch_phase_rhs_fe<2, AMPE_FLUX_SIMPLE>(L, p, fe, phi, q, cl, cs, rhs, s);
   }
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_quat_rhs(const Level& L, const ampe_model_params& p, const double* phi,
                    const double* q, const double* n2, double* rhs, hipStream_t s)
{
   RhsConst k = rhs_const(p);
   dim3 g(L.ncol_tiles), b = tile_block();
   if (L.ndim == 3)
      hipLaunchKernelGGL((k_quat_rhs<3, false>), g, b, 0, s, level_args_col(L), p, k, phi, q, n2,
                         rhs);
   else
      hipLaunchKernelGGL((k_quat_rhs<2, false>), g, b, 0, s, level_args_col(L), p, k, phi, q, n2,
                         rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_quat_rhs_fc(const Level& L, const ampe_model_params& p, const double* phi,
                       const double* q, const double* fc, double* rhs, hipStream_t s)
{
   RhsConst k = rhs_const(p);
   dim3 g(L.ncol_tiles), b = tile_block();
   if (L.ndim == 3)
      hipLaunchKernelGGL((k_quat_rhs<3, true>), g, b, 0, s, level_args_col(L), p, k, phi, q, fc,
                         rhs);
   else
      hipLaunchKernelGGL((k_quat_rhs<2, true>), g, b, 0, s, level_args_col(L), p, k, phi, q, fc,
                         rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_quat_project(const Level& L, const ampe_model_params& p, const double* q,
                        double* corr, double* err, hipStream_t s)
{
   hipLaunchKernelGGL(k_quat_project, dim3(L.ncell_tiles), tile_block(), 0, s, level_args(L),
                      p.qlen, q, corr, err);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_conc_coeffs(const Level& L, const ampe_model_params& p, const double* phi,
                       const double* cl, const double* cs, double* Dc, double* Dphi,
                       hipStream_t s)
{
   AMPE_DISPATCH(L, k_conc_coeffs, L.nface_tiles, level_args(L, true), p, rhs_const(p), phi, cl, cs, Dc, Dphi);
   return 0;
}

int launch_conc_rhs(const Level& L, const ampe_model_params& p, const double* c,
                    const double* phi, const double* Dc, const double* Dphi, double* rhs,
                    hipStream_t s)
{
   AMPE_DISPATCH(L, k_conc_rhs, L.ncol_tiles, level_args_col(L), p, rhs_const(p), c, phi, Dc, Dphi, rhs);
   return 0;
}

int launch_ebs_coeffs(const Level& L, const ampe_model_params& p, const double* phi, double* DL,
                      double* DA, hipStream_t s)
{
   AMPE_DISPATCH(L, k_ebs_coeffs, L.nface_tiles, level_args(L, true), p, rhs_const(p), phi, DL, DA);
   return 0;
}

int launch_ebs_coeffs_mob(const Level& L, const ampe_model_params& p, const double* phi,
                          const double* cl, const double* cs, double* DL, double* DA,
                          hipStream_t s)
{
   AMPE_DISPATCH(L, k_ebs_coeffs_mob, L.nface_tiles, level_args(L, true), p, mob_const(p), phi,
                 cl, cs, DL, DA);
   return 0;
}

int launch_conc_rhs_ebs(const Level& L, const ampe_model_params& p, const double* cl,
                        const double* cs, const double* DL, const double* DA, double* rhs,
                        hipStream_t s)
{
   AMPE_DISPATCH(L, k_conc_rhs_ebs, L.ncol_tiles, level_args_col(L), p, rhs_const(p), cl, cs, DL, DA, rhs);
   return 0;
}

int launch_conc_rhs_ebs_4th(const Level& L, const ampe_model_params& p, const double* cl,
                            const double* cs, const double* DL, const double* DA, double* rhs,
                            hipStream_t s)
{
   if (L.g < 2) {
      set_error("conc_rhs_ebs_4th: the level needs two ghost cells");
      return -1;
   }
   AMPE_DISPATCH(L, k_conc_rhs_ebs_4th, L.ncol_tiles, level_args_col(L), p, rhs_const(p), cl, cs,
                 DL, DA, rhs);
   return 0;
}

int launch_conc_rhs_ebs_iso(const Level& L, const ampe_model_params& p, const double* phid,
                            const double* cl, const double* cs, double* rhs, hipStream_t s)
{
   AMPE_DISPATCH(L, k_conc_rhs_ebs_iso, L.ncol_tiles, level_args_col(L), p, rhs_const(p), phid,
                 cl, cs, rhs);
   return 0;
}

int launch_tag(const Level& L, const ampe_model_params& p, const double* phi, double tol,
               int* tags, hipStream_t s)
{
   AMPE_DISPATCH(L, k_tag, L.ncell_tiles, level_args(L), p, rhs_const(p), phi, tol * tol, tags);
   return 0;
}

}  // namespace ampe

namespace ampe {

// Face coefficients of the orientation operator on every face (used by the
// preconditioner; QuatFACOps::setOperatorCoefficients,
// source/QuatFACOps.cc:735-829).
template <int NDIM>
__global__ void __launch_bounds__(256) k_quat_face_coef(LevelArgs a, ampe_model_params p, RhsConst k,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ n2,
                                                        double* __restrict__ fc)
{
   CellPos c;
   if (!cell_pos(a, 1, c)) return;
   if (NDIM == 2 && c.k > 0) return;
   for (int d = 0; d < NDIM; d++) {
      bool ok = (c.i < c.pd.nx + (d == 0)) && (c.j < c.pd.ny + (d == 1)) &&
                (NDIM == 2 || c.k < c.pd.nz + (d == 2));
      if (!ok) continue;
      long long fi = face_idx(a, c.pd, d, c.i, c.j, c.k);
      fc[fi] = quat_face_coef(p, k, phi[c.c0 - c.st[d]], phi[c.c0], n2[fi]);
   }
}

// sqrt of the orientation mobility over all ghosted cells
__global__ void __launch_bounds__(256) k_sqrt_quat_mobility(long long n, ampe_model_params p,
                                                            const double* __restrict__ phi,
                                                            double* __restrict__ sm)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   sm[t] = sqrt(quat_mobility(p, phi[t]));
}

int launch_quat_face_coef(const Level& L, const ampe_model_params& p, const double* phi,
                          const double* n2, double* fc, hipStream_t s)
{
   AMPE_DISPATCH(L, k_quat_face_coef, L.nface_tiles, level_args(L, true), p, rhs_const(p), phi, n2, fc);
   return 0;
}

int launch_sqrt_quat_mobility(const Level& L, const ampe_model_params& p, const double* phi,
                              double* sm, hipStream_t s)
{
   hipLaunchKernelGGL(k_sqrt_quat_mobility, dim3((unsigned)((L.ncg + 255) / 256)), dim3(256), 0,
                      s, L.ncg, p, phi, sm);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

// ------------------------------------------------------------ hierarchy
// Composite fluxes of PhaseRHSStrategyWithQ::evaluateRHS (source/
// PhaseRHSStrategyWithQ.cc:126-144), QuatIntegrator::
// evaluateConcentrationRHS (source/QuatIntegrator.cc:2610-2620) and
// QuatFACOps::accumulateOperatorOnLevel (source/QuatFACOps.cc:2040-2100):
// the coarse face on the coarse-fine boundary carries the side-data
// CONSERVATIVE_COARSEN of the fine fluxes (sum of fine flux * fine face area,
// divided by the coarse face area).  The fused coarse kernel used the coarse
// flux there; one thread per listed coarse cell replaces it in the cell's
// divergence, with the same face-flux functions as the fused kernels.
template <int NDIM>
__global__ void __launch_bounds__(256) k_hier_phase_corr(const CfCorr* __restrict__ cc, int n,
                                                         const CfFace* __restrict__ ff,
                                                         ampe_model_params pc, RhsConst kc,
                                                         ampe_model_params pf, RhsConst kf,
                                                         long long ncg,
                                                         const double* __restrict__ phi,
                                                         const double* __restrict__ q,
                                                         const double* __restrict__ cl,
                                                         const double* __restrict__ cs, double af,
                                                         double ac, double* __restrict__ rhs)
{
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfCorr& e = cc[t];
   const long long stc[3] = {1, e.sy, e.sz};
   const int nf = NDIM == 3 ? 4 : 2;
   double delta = 0.0;
   for (int f = 0; f < e.nface; f++) {
      const int d = e.d[f];
      double fold = phase_face_flux<NDIM>(pc, kc, phi, q, ncg, stc, d, e.cm[f], e.c0[f]);
      double acc = 0.0;
      for (int r = 0; r < nf; r++) {
         const CfFace& g = ff[e.fine[f] + r];
         const long long stf[3] = {1, g.sy, g.sz};
         acc = acc + phase_face_flux<NDIM>(pf, kf, phi, q, ncg, stf, d, g.cm, g.c0) * af;
      }
      double fnew = acc / ac;
      double t1 = (fnew - fold) * kc.dxinv[d];
      delta = delta + (e.side[f] ? t1 : -t1);
   }
   double mob = pc.phi_mobility;
   if (pc.with_conc && pc.phi_mobility_type == AMPE_PHI_MOB_KIM)
      mob = phase_mobility(pc, cl[e.cell], cs[e.cell]);
   rhs[e.cell] = rhs[e.cell] + mob * delta;
}

template <int NDIM>
__global__ void __launch_bounds__(256) k_hier_conc_corr(const CfCorr* __restrict__ cc, int n,
                                                        const CfFace* __restrict__ ff,
                                                        ampe_model_params pc, RhsConst kc,
                                                        RhsConst kf,
                                                        const double* __restrict__ c,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ Dc,
                                                        const double* __restrict__ Dphi,
                                                        double af, double ac,
                                                        double* __restrict__ rhs)
{
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfCorr& e = cc[t];
   const int nf = NDIM == 3 ? 4 : 2;
   double delta = 0.0;
   for (int f = 0; f < e.nface; f++) {
      const int d = e.d[f];
      const long long fo = e.face[f], m0 = e.cm[f], p0 = e.c0[f];
      double fold = kc.dxinv[d] * (Dc[fo] * (c[p0] - c[m0]) + Dphi[fo] * (phi[p0] - phi[m0]));
      double acc = 0.0;
      for (int r = 0; r < nf; r++) {
         const CfFace& g = ff[e.fine[f] + r];
         double jf = kf.dxinv[d] * (Dc[g.face] * (c[g.c0] - c[g.cm]) +
                                    Dphi[g.face] * (phi[g.c0] - phi[g.cm]));
         acc = acc + jf * af;
      }
      double fnew = acc / ac;
      double t1 = kc.dxinv[d] * (fnew - fold);
      delta = delta + (e.side[f] ? t1 : -t1);
   }
   rhs[e.cell] = rhs[e.cell] + pc.conc_mobility * delta;
}

template <int NDIM>
__global__ void __launch_bounds__(256) k_hier_quat_corr(const CfCorr* __restrict__ cc, int n,
                                                        const CfFace* __restrict__ ff,
                                                        ampe_model_params pc, RhsConst kc,
                                                        RhsConst kf, long long ncg,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ q,
                                                        const double* __restrict__ fc,
                                                        double af, double ac,
                                                        double* __restrict__ rhs)
{
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfCorr& e = cc[t];
   const int nf = NDIM == 3 ? 4 : 2;
   const int ql = pc.qlen;
   const long long c0 = e.cell;
   double ddv[4] = {0.0, 0.0, 0.0, 0.0}, dlam = 0.0, sumq2 = 0.0;
   for (int m = 0; m < ql; m++) {
      const double* aa = q + m * ncg;
      double qm = aa[c0];
      double dl = 0.0;
      for (int f = 0; f < e.nface; f++) {
         const int d = e.d[f];
         double fold = fc[e.face[f]] * (kc.dxinv[d] * (aa[e.c0[f]] - aa[e.cm[f]]));
         double acc = 0.0;
         for (int r = 0; r < nf; r++) {
            const CfFace& g = ff[e.fine[f] + r];
            acc = acc + fc[g.face] * (kf.dxinv[d] * (aa[g.c0] - aa[g.cm])) * af;
         }
         double fnew = acc / ac;
         double dif = e.side[f] ? (fnew - fold) : -(fnew - fold);
         ddv[m] = ddv[m] + dif * kc.dxinv[d];
         dl = dl + dif * kc.hdxinv[d];
      }
      dlam = dlam - qm * dl;
      sumq2 = sumq2 + qm * qm;
   }
   dlam = dlam / sumq2;
   double mq = quat_mobility(pc, phi[c0]);
   for (int m = 0; m < ql; m++) {
      double qm = q[m * ncg + c0];
      rhs[m * ncg + c0] = rhs[m * ncg + c0] - mq * (ddv[m] + 2.0 * qm * dlam);
   }
}

int launch_hier_phase_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                           const ampe_model_params& pc, const ampe_model_params& pf,
                           long long ncg, const double* phi, const double* q, const double* cl,
                           const double* cs, double af, double ac, double* rhs, hipStream_t s)
{
   if (ncorr <= 0) return 0;
   RhsConst kc = rhs_const(pc), kf = rhs_const(pf);
   dim3 g((ncorr + 255) / 256), b(256);
   if (ndim == 3)
      hipLaunchKernelGGL(k_hier_phase_corr<3>, g, b, 0, s, cc, ncorr, ff, pc, kc, pf, kf, ncg,
                         phi, q, cl, cs, af, ac, rhs);
   else
      hipLaunchKernelGGL(k_hier_phase_corr<2>, g, b, 0, s, cc, ncorr, ff, pc, kc, pf, kf, ncg,
                         phi, q, cl, cs, af, ac, rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_hier_conc_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                          const ampe_model_params& pc, const ampe_model_params& pf,
                          const double* c, const double* phi, const double* Dc,
                          const double* Dphi, double af, double ac, double* rhs, hipStream_t s)
{
   if (ncorr <= 0) return 0;
   RhsConst kc = rhs_const(pc), kf = rhs_const(pf);
   dim3 g((ncorr + 255) / 256), b(256);
   if (ndim == 3)
      hipLaunchKernelGGL(k_hier_conc_corr<3>, g, b, 0, s, cc, ncorr, ff, pc, kc, kf, c, phi, Dc,
                         Dphi, af, ac, rhs);
   else
      hipLaunchKernelGGL(k_hier_conc_corr<2>, g, b, 0, s, cc, ncorr, ff, pc, kc, kf, c, phi, Dc,
                         Dphi, af, ac, rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_hier_quat_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                          const ampe_model_params& pc, const ampe_model_params& pf,
                          long long ncg, const double* phi, const double* q, const double* fc,
                          double af, double ac, double* rhs, hipStream_t s)
{
   if (ncorr <= 0) return 0;
   if (pc.qlen > 4) {
      set_error("launch_hier_quat_corr: qlen must be <= 4");
      return -1;
   }
   RhsConst kc = rhs_const(pc), kf = rhs_const(pf);
   dim3 g((ncorr + 255) / 256), b(256);
   if (ndim == 3)
      hipLaunchKernelGGL(k_hier_quat_corr<3>, g, b, 0, s, cc, ncorr, ff, pc, kc, kf, ncg, phi, q,
                         fc, af, ac, rhs);
   else
      hipLaunchKernelGGL(k_hier_quat_corr<2>, g, b, 0, s, cc, ncorr, ff, pc, kc, kf, ncg, phi, q,
                         fc, af, ac, rhs);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

}  // namespace ampe
