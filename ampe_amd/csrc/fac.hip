// FAC preconditioner kernels: phase operator coefficient C, red-black
// Gauss-Seidel colour sweeps and residuals of A u = M div(D grad u) + C u, and
// the orientation block's symmetrised operator I + gamma sqrt(M) L sqrt(M).
// Templated on the storage/compute type: double for the FP64 path, float for
// the mixed-precision preconditioner.
#include <type_traits>

#include "fac.h"

namespace ampe {

// PhaseFACOps::setCOnPatchPrivate (source/PhaseFACOps.cc:101-183)
__global__ void __launch_bounds__(256) k_phase_C(LevelArgs a, ampe_model_params p, double gamma,
                                                 const double* __restrict__ phi,
                                                 const double* __restrict__ M,
                                                 double* __restrict__ C)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   double g2 = d2well(phi[c.c0], p.phi_well_type);
   double gm = gamma * (M ? M[c.c0] : p.phi_mobility);
   C[c.c0] = 1.0 + gm * p.phi_well_scale * g2;
}

// Colour of the update: efo_rbgswithfluxmax* walk row (j, k) from
// ifirst + ioffset in steps of 2, with ioffset = 1 when the parity of
// (jfirst + j) [2-D] or (ifirst + j + k) [3-D] differs from the colour
// (source/fortran/2d/ellipticfacops.m4:111-130, 3d/ellipticfacops.m4).
template <int NDIM>
__device__ __forceinline__ bool is_color(const CellPos& c, int color)
{
   int gj = c.pd.lo[1] + c.j;
   int x = (NDIM == 3) ? c.pd.lo[0] + gj + (c.pd.lo[2] + c.k) : c.pd.lo[1] + gj;
   int io = ((x - (x / 2) * 2) != color) ? 1 : 0;
   return ((c.i - io) & 1) == 0 && c.i >= io;
}

// Colour-compact mapping: a workgroup of TX/2 x TY threads covers one tile,
// each thread the one cell of the colour in its pair (2 tx, 2 tx + 1) of the
// row, so every lane updates a cell (the full-tile mapping leaves half of
// them idle).  Plane t.k + kk; false when outside the box.
template <int NDIM>
__device__ __forceinline__ bool cell_pos_color(const LevelArgs& a, int kk, int color,
                                               CellPos& c)
{
   const Tile t = a.T[tile_index(a)];
   c.pd = a.P[t.patch];
   c.j = t.j0 + (int)threadIdx.y;
   c.k = t.k + kk;
   int gj = c.pd.lo[1] + c.j;
   int x = (NDIM == 3) ? c.pd.lo[0] + gj + (c.pd.lo[2] + c.k) : c.pd.lo[1] + gj;
   int io = ((x - (x / 2) * 2) != color) ? 1 : 0;
   c.i = t.i0 + 2 * (int)threadIdx.x + io;  // t.i0 is a multiple of TX (even)
   if (c.i >= c.pd.nx || c.j >= c.pd.ny || c.k >= c.pd.nz) return false;
   c.st[0] = 1;
   c.st[1] = c.pd.sy;
   c.st[2] = c.pd.sz;
   c.c0 = c.pd.cell_off + (c.i + a.g) + (long long)(c.j + a.g) * c.pd.sy +
          (long long)(c.k + a.gz) * c.pd.sz;
   return true;
}

template <typename T>
struct OpArgs {
   const T* D;
   T Dc;
   const T* C;
   T Cc;
   const T* M;
   T Mc;
};

// res = rhs - M div(D grad u) - C u at one cell; *diag = M sum D/h^2 - C
// (the reciprocal is dudr of efo_rbgswithfluxmax*).
template <typename T, int NDIM>
__device__ __forceinline__ T op_residual(const LevelArgs& a, const CellPos& c, const OpArgs<T>& op,
                                         const T* __restrict__ rhs, const T* __restrict__ u,
                                         const T* dinv, T* diag)
{
   const long long c0 = c.c0;
   const int ijk[3] = {c.i, c.j, c.k};
   T div = 0, dsum = 0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      T dlo, dhi;
      if (op.D) {
         int lo[3] = {ijk[0], ijk[1], ijk[2]};
         long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
         lo[d] += 1;
         long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
         dlo = op.D[flo];
         dhi = op.D[fhi];
      } else {
         dlo = dhi = op.Dc;
      }
      long long s = c.st[d];
      T flo = dinv[d] * dlo * (u[c0] - u[c0 - s]);
      T fhi = dinv[d] * dhi * (u[c0 + s] - u[c0]);
      T t = dinv[d] * (fhi - flo);
      T ds = dinv[d] * dinv[d] * (dhi + dlo);
      div = (d == 0) ? t : div + t;
      dsum = (d == 0) ? ds : dsum + ds;
   }
   T m = op.M ? op.M[c0] : op.Mc;
   T cc = op.C ? op.C[c0] : op.Cc;
   *diag = m * dsum - cc;
   return rhs[c0] - m * div - cc * u[c0];
}

template <typename T>
__device__ __forceinline__ void block_store_max(T v, double* partial)
{
   __shared__ double sh[TX * TY];
   int t = threadIdx.x + TX * threadIdx.y;
   sh[t] = (double)v;
   __syncthreads();
   for (int w = TX * TY / 2; w > 0; w >>= 1) {
      if (t < w) sh[t] = sh[t + w] > sh[t] ? sh[t + w] : sh[t];
      __syncthreads();
   }
   if (t == 0) partial[blockIdx.x] = sh[0];
}

template <typename T>
__device__ __forceinline__ void block_store_sum(double v, double* partial)
{
   __shared__ double sh[TX * TY];
   int t = threadIdx.x + TX * threadIdx.y;
   sh[t] = v;
   __syncthreads();
   for (int w = TX * TY / 2; w > 0; w >>= 1) {
      if (t < w) sh[t] = sh[t] + sh[t + w];
      __syncthreads();
   }
   if (t == 0) partial[blockIdx.x] = sh[0];
}

// One colour of efo_rbgswithfluxmax{var,con}dc{var,con}sf
// (source/fortran/2d/ellipticfacops.m4:60-345): u += -res * dudr.
template <typename T, int NDIM>
__device__ __forceinline__ T rbgs_cell(const LevelArgs& a, const CellPos& c, const OpArgs<T>& op,
                                       const T* dinv, const T* __restrict__ rhs, T* __restrict__ u)
{
   T diag;
   T res = op_residual<T, NDIM>(a, c, op, rhs, u, dinv, &diag);
   T dudr = T(1) / diag;
   T du = -res * dudr;
   u[c.c0] = u[c.c0] + du * T(1);
   return res < 0 ? -res : res;
}

template <typename T, int NDIM>
__global__ void __launch_bounds__(256) k_rbgs(LevelArgs a, OpArgs<T> op, T di0, T di1, T di2,
                                              const T* __restrict__ rhs, T* __restrict__ u,
                                              int color, double* __restrict__ partial)
{
   T absres = 0;
   const T dinv[3] = {di0, di1, di2};
   AMPE_COLUMN_BEGIN(a, c)
   if (!is_color<NDIM>(c, color)) continue;
   T ar = rbgs_cell<T, NDIM>(a, c, op, dinv, rhs, u);
   absres = ar > absres ? ar : absres;
   AMPE_COLUMN_END
   if (partial) block_store_max<T>(absres, partial);
}

// The same on the colour-compact mapping (no max-residual partials).
template <typename T, int NDIM>
__global__ void __launch_bounds__(128) k_rbgs_c(LevelArgs a, OpArgs<T> op, T di0, T di1, T di2,
                                                const T* __restrict__ rhs, T* __restrict__ u,
                                                int color)
{
   const T dinv[3] = {di0, di1, di2};
   for (int kk = 0; kk < a.nz_tile; kk++) {
      CellPos c;
      if (cell_pos_color<NDIM>(a, kk, color, c)) (void)rbgs_cell<T, NDIM>(a, c, op, dinv, rhs, u);
   }
}

// efo_compres*sca (source/fortran/2d/ellipticfacops.m4:346-446) + sum res^2
template <typename T, int NDIM>
__global__ void __launch_bounds__(256) k_residual(LevelArgs a, OpArgs<T> op, T di0, T di1, T di2,
                                                  const T* __restrict__ rhs, const T* __restrict__ u,
                                                  T* __restrict__ res, double* __restrict__ partial)
{
   double sq = 0.0;
   AMPE_COLUMN_BEGIN(a, c)
   T dinv[3] = {di0, di1, di2};
   T diag;
   T r = op_residual<T, NDIM>(a, c, op, rhs, u, dinv, &diag);
   res[c.c0] = r;
   sq += (double)r * (double)r;
   AMPE_COLUMN_END
   if (partial) block_store_sum<T>(sq, partial);
}

// The elliptic residual restricted from the red cells only: after red-black
// sweeps ending with the black colour each black row's residual vanishes (its
// update zeroes it, and the row couples the cell to red neighbours only), so
// the coarse right-hand side is, per coarse cell, the mean of its red
// children's residuals in k_restrict_cell's summation order.  Periodic levels
// with one checkerboard (the caller checks).
template <typename T, int NDIM>
__global__ void __launch_bounds__(256) k_ell_resrestrict(LevelArgs ca, LevelArgs fa,
                                                         OpArgs<T> op, T di0, T di1, T di2,
                                                         const T* __restrict__ rhs,
                                                         const T* __restrict__ u,
                                                         T* __restrict__ b)
{
   CellPos p;
   if (!cell_pos(ca, 0, p)) return;
   const T dinv[3] = {di0, di1, di2};
   const PatchDev fp = fa.P[ca.T[tile_index(ca)].patch];
   T pr[4];
#pragma unroll
   for (int dz = 0; dz < (NDIM == 3 ? 2 : 1); dz++)
#pragma unroll
      for (int dy = 0; dy < 2; dy++) {
         CellPos c;
         c.pd = fp;
         c.j = 2 * p.j + dy;
         c.k = NDIM == 3 ? 2 * p.k + dz : 0;
         int gj = fp.lo[1] + c.j;
         int x = (NDIM == 3) ? fp.lo[0] + gj + (fp.lo[2] + c.k) : fp.lo[1] + gj;
         int io = ((x - (x / 2) * 2) != 0) ? 1 : 0;  // red cells of the row (cell_pos_color)
         c.i = 2 * p.i + io;
         c.st[0] = 1;
         c.st[1] = fp.sy;
         c.st[2] = fp.sz;
         c.c0 = fp.cell_off + (c.i + fa.g) + (long long)(c.j + fa.g) * fp.sy +
                (long long)(c.k + fa.gz) * fp.sz;
         T diag;
         pr[dy + 2 * dz] = op_residual<T, NDIM>(fa, c, op, rhs, u, dinv, &diag);
      }
   T sum = pr[0] + pr[1];
   if (NDIM == 3) sum = sum + (pr[2] + pr[3]);
   b[p.c0] = sum * (NDIM == 3 ? T(0.125) : T(0.25));
}

// ------------------------------------------------------------ orientation
// Row of A = I + gamma sqrt(M) L sqrt(M), L from face coefficients:
// off-diagonal fc/h^2, diagonal -sum fc/h^2 (set_j_ij2d/3d, set_stencil2d/3d,
// source/fortran/2d/quatlevelsolver.m4:9-120).  Returns A z at one cell for
// component block zc and the centre coefficient.
template <typename T, int NDIM>
__device__ __forceinline__ T quat_row(const LevelArgs& a, const CellPos& c, const T* __restrict__ fc,
                                      const T* __restrict__ sm, T gamma, const T* hinv2,
                                      const T* __restrict__ z, T* center);

// The same row with the z values supplied by zv(d, side): the cell itself for
// d < 0, else its neighbour across the low (side -1) or high (+1) face of
// direction d.  Same operations in the same order as quat_row.
template <typename T, int NDIM, class ZV>
__device__ __forceinline__ T quat_row_zv(const LevelArgs& a, const CellPos& c,
                                         const T* __restrict__ fc, const T* __restrict__ sm,
                                         T gamma, const T* hinv2, const ZV& zv, T* center)
{
   const long long c0 = c.c0;
   const int ijk[3] = {c.i, c.j, c.k};
   T olo[3], ohi[3];
   T diag = 0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      int lo[3] = {ijk[0], ijk[1], ijk[2]};
      long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      lo[d] += 1;
      long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      olo[d] = fc[flo] * hinv2[d];
      ohi[d] = fc[fhi] * hinv2[d];
      diag = (d == 0) ? -olo[0] - ohi[0] : diag - olo[d] - ohi[d];
   }
   T fac = gamma * sm[c0];
   *center = fac * diag * sm[c0] + T(1);
   T r = *center * zv(-1, 0);
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      long long s = c.st[d];
      r = r + fac * olo[d] * sm[c0 - s] * zv(d, -1) + fac * ohi[d] * sm[c0 + s] * zv(d, 1);
   }
   return r;
}

template <typename T, int NDIM>
__device__ __forceinline__ T quat_row(const LevelArgs& a, const CellPos& c, const T* __restrict__ fc,
                                      const T* __restrict__ sm, T gamma, const T* hinv2,
                                      const T* __restrict__ z, T* center)
{
   const long long c0 = c.c0;
   const int ijk[3] = {c.i, c.j, c.k};
   T olo[3], ohi[3];
   T diag = 0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      int lo[3] = {ijk[0], ijk[1], ijk[2]};
      long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      lo[d] += 1;
      long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      olo[d] = fc[flo] * hinv2[d];
      ohi[d] = fc[fhi] * hinv2[d];
      diag = (d == 0) ? -olo[0] - ohi[0] : diag - olo[d] - ohi[d];
   }
   T fac = gamma * sm[c0];
   *center = fac * diag * sm[c0] + T(1);
   T r = *center * z[c0];
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      long long s = c.st[d];
      r = r + fac * olo[d] * sm[c0 - s] * z[c0 - s] + fac * ohi[d] * sm[c0 + s] * z[c0 + s];
   }
   return r;
}

// The row's coefficients at one cell, shared by every component: centre
// gamma sm^2 diag + 1 and the neighbour weights gamma sm fc/h^2 sm_nbr (the
// products quat_row forms, in its order).
template <typename T>
struct QuatRowCoef {
   T center;
   T wlo[3], whi[3];
};

template <typename T, int NDIM>
__device__ __forceinline__ QuatRowCoef<T> quat_row_coef(const LevelArgs& a, const CellPos& c,
                                                        const T* __restrict__ fc,
                                                        const T* __restrict__ sm, T gamma,
                                                        const T* hinv2)
{
   const long long c0 = c.c0;
   const int ijk[3] = {c.i, c.j, c.k};
   T olo[3], ohi[3];
   T diag = 0;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      int lo[3] = {ijk[0], ijk[1], ijk[2]};
      long long flo = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      lo[d] += 1;
      long long fhi = face_idx(a, c.pd, d, lo[0], lo[1], lo[2]);
      olo[d] = fc[flo] * hinv2[d];
      ohi[d] = fc[fhi] * hinv2[d];
      diag = (d == 0) ? -olo[0] - ohi[0] : diag - olo[d] - ohi[d];
   }
   QuatRowCoef<T> w;
   T fac = gamma * sm[c0];
   w.center = fac * diag * sm[c0] + T(1);
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      w.wlo[d] = fac * olo[d] * sm[c0 - c.st[d]];
      w.whi[d] = fac * ohi[d] * sm[c0 + c.st[d]];
   }
   return w;
}

// Component values of one cell's row: the cell and its face neighbours.
template <typename T, int NDIM>
struct QuatRowVals {
   T zc, zl[NDIM], zh[NDIM];
};

// Where one cell's component values sit: the cell and its two neighbours in
// the row (rows above and below, planes before and after: the cell's index
// -+ the row / plane stride).  SPLIT: the row parity-split layout.
struct QuatIdx {
   long long c, xl, xh;
};

template <bool SPLIT>
__device__ __forceinline__ QuatIdx quat_idx(const LevelArgs& a, const CellPos& c)
{
   if (SPLIT) return {split_at(a, c, 0), split_at(a, c, -1), split_at(a, c, 1)};
   return {c.c0, c.c0 - 1, c.c0 + 1};
}

template <typename T, int NDIM>
__device__ __forceinline__ void quat_row_load(const CellPos& c, const QuatIdx& q,
                                              const T* __restrict__ zm, QuatRowVals<T, NDIM>& v)
{
   v.zc = zm[q.c];
   v.zl[0] = zm[q.xl];
   v.zh[0] = zm[q.xh];
#pragma unroll
   for (int d = 1; d < NDIM; d++) {
      v.zl[d] = zm[q.c - c.st[d]];
      v.zh[d] = zm[q.c + c.st[d]];
   }
}

// A z from the loaded values: the sum quat_row forms, in its order.
template <typename T, int NDIM>
__device__ __forceinline__ T quat_row_apply(const QuatRowCoef<T>& w, const QuatRowVals<T, NDIM>& v)
{
   T r = w.center * v.zc;
#pragma unroll
   for (int d = 0; d < NDIM; d++) r = r + w.wlo[d] * v.zl[d] + w.whi[d] * v.zh[d];
   return r;
}

// Gauss-Seidel update of one cell, every component.  QL > 0: QL components,
// all loads issued before the first update (the updates write only this
// cell, which no other load of the cell reads); QL = 0: ncomp components one
// after the other.  SPLIT: z and rhs in the row parity-split layout (fc and
// sm natural).
template <typename T, int NDIM, int QL, bool SPLIT>
__device__ __forceinline__ void quat_rbgs_cell(const LevelArgs& a, const CellPos& c,
                                               const T* __restrict__ fc, const T* __restrict__ sm,
                                               T gamma, const T* hinv2, int ncomp,
                                               const T* __restrict__ rhs, T* __restrict__ z)
{
   const QuatRowCoef<T> w = quat_row_coef<T, NDIM>(a, c, fc, sm, gamma, hinv2);
   const QuatIdx q = quat_idx<SPLIT>(a, c);
   if (QL > 0) {
      QuatRowVals<T, NDIM> v[QL > 0 ? QL : 1];
      T rv[QL > 0 ? QL : 1];
#pragma unroll
      for (int m = 0; m < QL; m++) {
         quat_row_load<T, NDIM>(c, q, z + m * a.ncg, v[m]);
         rv[m] = rhs[m * a.ncg + q.c];
      }
#pragma unroll
      for (int m = 0; m < QL; m++) {
         T res = rv[m] - quat_row_apply<T, NDIM>(w, v[m]);
         z[m * a.ncg + q.c] = v[m].zc + res / w.center;
      }
   } else {
      for (int m = 0; m < ncomp; m++) {
         QuatRowVals<T, NDIM> v;
         quat_row_load<T, NDIM>(c, q, z + m * a.ncg, v);
         T res = rhs[m * a.ncg + q.c] - quat_row_apply<T, NDIM>(w, v);
         z[m * a.ncg + q.c] = v.zc + res / w.center;
      }
   }
}

// One colour of the sweep on the colour-compact mapping.
template <typename T, int NDIM, int QL, bool SPLIT>
__global__ void __launch_bounds__(128) k_quat_rbgs_c(LevelArgs a, const T* __restrict__ fc,
                                                     const T* __restrict__ sm, T gamma, T hi0,
                                                     T hi1, T hi2, int ncomp,
                                                     const T* __restrict__ rhs,
                                                     T* __restrict__ z, int color)
{
   const T hinv2[3] = {hi0, hi1, hi2};
   for (int kk = 0; kk < a.nz_tile; kk++) {
      CellPos c;
      if (cell_pos_color<NDIM>(a, kk, color, c))
         quat_rbgs_cell<T, NDIM, QL, SPLIT>(a, c, fc, sm, gamma, hinv2, ncomp, rhs, z);
   }
}


// ------------------------------------------------------------ fused sweep
// One red-black sweep of the orientation block in one pass, zin -> zout
// (zin ghost-filled, not written).  A workgroup first updates every red cell
// of its tile grown by one cell in each direction (neighbouring tiles'
// cells included, from zin only) into LDS, then the black cells of its tile
// from those red values: the same operations as the red launch followed by
// the black launch, so the same values, with zin read once instead of twice.
// Black cells on a patch's outer layer need values of ghost cells; they keep
// their previous value here and are updated by k_quat_rbgs_shell after a
// ghost fill of zout.
constexpr int kRX = TX + 2, kRY = TY + 2;

template <int NDIM>
__device__ __forceinline__ void make_cell(const LevelArgs& a, const PatchDev& pd, int i, int j,
                                          int k, CellPos& c)
{
   c.pd = pd;
   c.i = i;
   c.j = j;
   c.k = k;
   c.st[0] = 1;
   c.st[1] = pd.sy;
   c.st[2] = pd.sz;
   c.c0 = pd.cell_off + (i + a.g) + (long long)(j + a.g) * pd.sy + (long long)(k + a.gz) * pd.sz;
}

template <int NDIM>
__device__ __forceinline__ bool outer_layer(const CellPos& c)
{
   return c.i == 0 || c.i == c.pd.nx - 1 || c.j == 0 || c.j == c.pd.ny - 1 ||
          (NDIM == 3 && (c.k == 0 || c.k == c.pd.nz - 1));
}

template <typename T, int NDIM>
__global__ void __launch_bounds__(256) k_quat_rbgs_fused(LevelArgs a, const T* __restrict__ fc,
                                                         const T* __restrict__ sm, T gamma,
                                                         T hi0, T hi1, T hi2, int ncomp,
                                                         const T* __restrict__ rhs,
                                                         const T* __restrict__ zin,
                                                         T* __restrict__ zout)
{
   // a column of the tile's planes [t.k, kend): the red values of the tile
   // grown by one cell in x and y for planes k-1, k, k+1 live in a ring of
   // three LDS slots (slot = plane % 3); each step adds plane k+1, then
   // updates the black cells of plane k
   constexpr int NP = kRX * kRY;
   constexpr int RING = NDIM == 3 ? 3 : 1;
   __shared__ T red[RING][4][NP];
   const T hinv2[3] = {hi0, hi1, hi2};
   const Tile t = a.T[tile_index(a)];
   const PatchDev pd = a.P[t.patch];
   const int kend = NDIM == 3 ? min(t.k + a.nz_tile, pd.nz) : 1;
   const int tid = (int)threadIdx.x + TX * (int)threadIdx.y;
   auto red_plane = [&](int p) {
      if (p < 0 || p >= pd.nz) return;
      const int slot = NDIM == 3 ? p % 3 : 0;
      for (int r = tid; r < NP; r += TX * TY) {
         const int ri = r % kRX - 1, rj = r / kRX - 1;
         const int i = t.i0 + ri, j = t.j0 + rj;
         if (i < 0 || i >= pd.nx || j < 0 || j >= pd.ny) continue;
         CellPos c;
         make_cell<NDIM>(a, pd, i, j, p, c);
         if (!is_color<NDIM>(c, 0)) continue;
         const bool own = ri >= 0 && ri < TX && rj >= 0 && rj < TY && p >= t.k && p < kend;
         for (int m = 0; m < ncomp; m++) {
            const T* zm = zin + m * a.ncg;
            T center;
            T az = quat_row<T, NDIM>(a, c, fc, sm, gamma, hinv2, zm, &center);
            T res = rhs[m * a.ncg + c.c0] - az;
            T v = zm[c.c0] + res / center;
            red[slot][m][r] = v;
            if (own) zout[m * a.ncg + c.c0] = v;
         }
      }
   };
   if (NDIM == 3) red_plane(t.k - 1);
   red_plane(t.k);
   const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
   const int i = t.i0 + tx, j = t.j0 + ty;
   const int r0 = (tx + 1) + kRX * (ty + 1);
   for (int k = t.k; k < kend; k++) {
      if (NDIM == 3) red_plane(k + 1);
      __syncthreads();
      if (i < pd.nx && j < pd.ny) {
         CellPos c;
         make_cell<NDIM>(a, pd, i, j, k, c);
         if (!is_color<NDIM>(c, 0)) {
            if (outer_layer<NDIM>(c)) {
               // previous value for now: the ghost fill of zout then gives
               // every ghost cell the value the black launch would read (red
               // updated, black not; ghost sources lie on their patches' outer
               // layers, and in 2-D the colouring is per patch, so a neighbour
               // across a patch face can be black)
               for (int m = 0; m < ncomp; m++) zout[m * a.ncg + c.c0] = zin[m * a.ncg + c.c0];
            } else {
               const int s0 = NDIM == 3 ? k % 3 : 0;
               const int sm1 = NDIM == 3 ? (k + 2) % 3 : 0, sp1 = NDIM == 3 ? (k + 1) % 3 : 0;
               for (int m = 0; m < ncomp; m++) {
                  const T* zm = zin + m * a.ncg;
                  auto zv = [&](int d, int side) -> T {
                     if (d < 0) return zm[c.c0];
                     if (d == 0) return red[s0][m][r0 + side];
                     if (d == 1) return red[s0][m][r0 + side * kRX];
                     return red[side < 0 ? sm1 : sp1][m][r0];
                  };
                  T center;
                  T az = quat_row_zv<T, NDIM>(a, c, fc, sm, gamma, hinv2, zv, &center);
                  T res = rhs[m * a.ncg + c.c0] - az;
                  zout[m * a.ncg + c.c0] = zm[c.c0] + res / center;
               }
            }
         }
      }
      __syncthreads();
   }
}

// Black cells on the outer layer: neighbours from zout (ghost-filled: red
// cells updated, black cells still previous), own previous value from zin.
template <typename T, int NDIM>
__global__ void __launch_bounds__(256) k_quat_rbgs_shell(LevelArgs a, const T* __restrict__ fc,
                                                         const T* __restrict__ sm, T gamma,
                                                         T hi0, T hi1, T hi2, int ncomp,
                                                         const T* __restrict__ rhs,
                                                         const T* __restrict__ zin,
                                                         T* __restrict__ zout)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   if (is_color<NDIM>(c, 0) || !outer_layer<NDIM>(c)) return;
   const T hinv2[3] = {hi0, hi1, hi2};
   for (int m = 0; m < ncomp; m++) {
      const T* zm = zin + m * a.ncg;
      const T* zo = zout + m * a.ncg;
      auto zv = [&](int d, int side) -> T {
         return d < 0 ? zm[c.c0] : zo[c.c0 + side * c.st[d]];
      };
      T center;
      T az = quat_row_zv<T, NDIM>(a, c, fc, sm, gamma, hinv2, zv, &center);
      T res = rhs[m * a.ncg + c.c0] - az;
      zout[m * a.ncg + c.c0] = zm[c.c0] + res / center;
   }
}

// Residual of one cell, every component (QL as in quat_rbgs_cell); returns
// the sum of the squared residuals (double).
template <typename T, int NDIM, int QL, bool SPLIT>
__device__ __forceinline__ double quat_residual_cell(const LevelArgs& a, const CellPos& c,
                                                     const T* __restrict__ fc,
                                                     const T* __restrict__ sm, T gamma,
                                                     const T* hinv2, int ncomp,
                                                     const T* __restrict__ rhs,
                                                     const T* __restrict__ z,
                                                     T* __restrict__ res)
{
   double sq = 0.0;
   const QuatRowCoef<T> w = quat_row_coef<T, NDIM>(a, c, fc, sm, gamma, hinv2);
   const QuatIdx q = quat_idx<SPLIT>(a, c);
   if (QL > 0) {
      QuatRowVals<T, NDIM> v[QL > 0 ? QL : 1];
      T rv[QL > 0 ? QL : 1];
#pragma unroll
      for (int m = 0; m < QL; m++) {
         quat_row_load<T, NDIM>(c, q, z + m * a.ncg, v[m]);
         rv[m] = rhs[m * a.ncg + q.c];
      }
#pragma unroll
      for (int m = 0; m < QL; m++) {
         T r = rv[m] - quat_row_apply<T, NDIM>(w, v[m]);
         res[m * a.ncg + q.c] = r;
         sq += (double)r * (double)r;
      }
   } else {
      for (int m = 0; m < ncomp; m++) {
         QuatRowVals<T, NDIM> v;
         quat_row_load<T, NDIM>(c, q, z + m * a.ncg, v);
         T r = rhs[m * a.ncg + q.c] - quat_row_apply<T, NDIM>(w, v);
         res[m * a.ncg + q.c] = r;
         sq += (double)r * (double)r;
      }
   }
   return sq;
}

template <typename T, int NDIM, int QL, bool SPLIT>
__global__ void __launch_bounds__(256) k_quat_residual(LevelArgs a, const T* __restrict__ fc,
                                                       const T* __restrict__ sm, T gamma, T hi0,
                                                       T hi1, T hi2, int ncomp,
                                                       const T* __restrict__ rhs,
                                                       const T* __restrict__ z, T* __restrict__ res,
                                                       double* __restrict__ partial)
{
   double sq = 0.0;
   const T hinv2[3] = {hi0, hi1, hi2};
   AMPE_COLUMN_BEGIN(a, c)
   sq += quat_residual_cell<T, NDIM, QL, SPLIT>(a, c, fc, sm, gamma, hinv2, ncomp, rhs, z, res);
   AMPE_COLUMN_END
   if (partial) block_store_sum<T>(sq, partial);
}

// The residual in the split layout without the reduction, on the pair
// mapping: a half-wave's cells have one column parity, so its split-layout
// loads and stores cover contiguous runs.
template <typename T, int NDIM, int QL>
__global__ void __launch_bounds__(128) k_quat_residual_p(LevelArgs a, const T* __restrict__ fc,
                                                         const T* __restrict__ sm, T gamma,
                                                         T hi0, T hi1, T hi2, int ncomp,
                                                         const T* __restrict__ rhs,
                                                         const T* __restrict__ z,
                                                         T* __restrict__ res)
{
   const T hinv2[3] = {hi0, hi1, hi2};
   for (int kk = 0; kk < a.nz_tile; kk++)
      for (int io = 0; io < 2; io++) {
         CellPos c;
         if (cell_pos_pair(a, kk, io, c))
            (void)quat_residual_cell<T, NDIM, QL, true>(a, c, fc, sm, gamma, hinv2, ncomp, rhs,
                                                        z, res);
      }
}

// Red residuals restricted.  After a pre-smoothing that ends with the black
// colour, the residual vanishes on black cells in exact arithmetic (a black
// update zeroes its own row's residual, and a black row couples its cell to
// red cells only), so the coarse right-hand side is the restriction of the red
// residuals: per coarse cell, its 2^(NDIM-1) red children's residuals in the
// restriction's summation order (each pair of x-children sums to its red
// member), without storing the fine residual.  Needs one checkerboard across
// patches and periodic ghosts (a zero-slope ghost would copy a black cell's
// new value into its own row); the host checks both.
template <typename T, int NDIM, int QL, bool SPLIT>
__global__ void __launch_bounds__(256) k_quat_resrestrict(LevelArgs ca, LevelArgs fa,
                                                          const T* __restrict__ fc,
                                                          const T* __restrict__ sm, T gamma,
                                                          T hi0, T hi1, T hi2,
                                                          const T* __restrict__ rhs,
                                                          const T* __restrict__ z,
                                                          T* __restrict__ b)
{
   CellPos p;
   if (!cell_pos(ca, 0, p)) return;
   const T hinv2[3] = {hi0, hi1, hi2};
   const PatchDev fp = fa.P[ca.T[tile_index(ca)].patch];
   T pr[QL][4];
#pragma unroll
   for (int dz = 0; dz < (NDIM == 3 ? 2 : 1); dz++)
#pragma unroll
      for (int dy = 0; dy < 2; dy++) {
         CellPos c;
         c.pd = fp;
         c.j = 2 * p.j + dy;
         c.k = NDIM == 3 ? 2 * p.k + dz : 0;
         int gj = fp.lo[1] + c.j;
         int x = (NDIM == 3) ? fp.lo[0] + gj + (fp.lo[2] + c.k) : fp.lo[1] + gj;
         int io = ((x - (x / 2) * 2) != 0) ? 1 : 0;  // red cells of the row (cell_pos_color)
         c.i = 2 * p.i + io;
         c.st[0] = 1;
         c.st[1] = fp.sy;
         c.st[2] = fp.sz;
         c.c0 = fp.cell_off + (c.i + fa.g) + (long long)(c.j + fa.g) * fp.sy +
                (long long)(c.k + fa.gz) * fp.sz;
         const QuatRowCoef<T> w = quat_row_coef<T, NDIM>(fa, c, fc, sm, gamma, hinv2);
         const QuatIdx q = quat_idx<SPLIT>(fa, c);
         QuatRowVals<T, NDIM> v[QL];
         T rv[QL];
#pragma unroll
         for (int m = 0; m < QL; m++) {
            quat_row_load<T, NDIM>(c, q, z + m * fa.ncg, v[m]);
            rv[m] = rhs[m * fa.ncg + q.c];
         }
#pragma unroll
         for (int m = 0; m < QL; m++) pr[m][dy + 2 * dz] = rv[m] - quat_row_apply<T, NDIM>(w, v[m]);
      }
   const long long co = SPLIT ? split_at(ca, p, 0) : p.c0;
#pragma unroll
   for (int m = 0; m < QL; m++) {
      T s = pr[m][0] + pr[m][1];
      if (NDIM == 3) s = s + (pr[m][2] + pr[m][3]);
      b[m * ca.ncg + co] = s * (NDIM == 3 ? T(0.125) : T(0.25));
   }
}

// ------------------------------------------------------------ launchers
static dim3 tblock() { return dim3(TX, TY); }
static dim3 cblock() { return dim3(TX / 2, TY); }

// Calls f(ndim, ql) with both as std::integral_constant: ql = ncomp for one,
// two or four components (all loads of a cell issued together), else 0.
template <class F>
static void quat_dispatch(int ndim, int ncomp, F&& f)
{
   using I2 = std::integral_constant<int, 2>;
   using I3 = std::integral_constant<int, 3>;
   auto by_ql = [&](auto nd) {
      switch (ncomp) {
      case 1: f(nd, std::integral_constant<int, 1>()); break;
      case 2: f(nd, std::integral_constant<int, 2>()); break;
      case 4: f(nd, std::integral_constant<int, 4>()); break;
      default: f(nd, std::integral_constant<int, 0>()); break;
      }
   };
   if (ndim == 3)
      by_ql(I3());
   else
      by_ql(I2());
}

// 1/h and 1/h^2 in the operator's precision, computed once per launch on the
// host (IEEE division: the device's value).
template <typename T>
static void recip(const double* dx, T* di)
{
   for (int d = 0; d < 3; d++) di[d] = T(1) / (T)dx[d];
}
template <typename T>
static void recip2(const double* dx, T* di)
{
   for (int d = 0; d < 3; d++) {
      T h = (T)dx[d];
      di[d] = T(1) / (h * h);
   }
}

int launch_phase_C(const Level& L, const ampe_model_params& p, double gamma,
                   const double* phi, const double* M, double* C, hipStream_t s)
{
   hipLaunchKernelGGL(k_phase_C, dim3(L.ncell_tiles), tblock(), 0, s, level_args(L), p, gamma,
                      phi, M, C);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

// Launch shape: column tiles (KZ planes per workgroup), or one-plane tiles
// when a per-tile reduction is requested (its callers reduce ncell_tiles
// partials).
static void shape(const Level& L, bool reduce, LevelArgs* a, int* grid)
{
   *a = reduce ? level_args(L) : level_args_col(L);
   *grid = reduce ? L.ncell_tiles : L.ncol_tiles;
}

template <typename T>
int launch_rbgs(const Level& L, const EllOp<T>& op, const double* dx, const T* rhs, T* u,
                int color, bool want_max, hipStream_t s)
{
   OpArgs<T> oa{op.D, op.Dc, op.C, op.Cc, op.M, op.Mc};
   double* part = want_max ? L.partial : nullptr;
   LevelArgs la;
   int grid;
   shape(L, want_max, &la, &grid);
   T di[3];
   recip(dx, di);
   if (!want_max) {
      if (L.ndim == 3)
         hipLaunchKernelGGL((k_rbgs_c<T, 3>), dim3(grid), cblock(), 0, s, la, oa, di[0], di[1],
                            di[2], rhs, u, color);
      else
         hipLaunchKernelGGL((k_rbgs_c<T, 2>), dim3(grid), cblock(), 0, s, la, oa, di[0], di[1],
                            di[2], rhs, u, color);
   } else if (L.ndim == 3) {
      hipLaunchKernelGGL((k_rbgs<T, 3>), dim3(grid), tblock(), 0, s, la, oa, di[0], di[1], di[2],
                         rhs, u, color, part);
   } else {
      hipLaunchKernelGGL((k_rbgs<T, 2>), dim3(grid), tblock(), 0, s, la, oa, di[0], di[1], di[2],
                         rhs, u, color, part);
   }
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int launch_residual(const Level& L, const EllOp<T>& op, const double* dx, const T* rhs,
                    const T* u, T* res, double* part, hipStream_t s)
{
   OpArgs<T> oa{op.D, op.Dc, op.C, op.Cc, op.M, op.Mc};
   LevelArgs la;
   int grid;
   shape(L, part != nullptr, &la, &grid);
   T di[3];
   recip(dx, di);
   if (L.ndim == 3)
      hipLaunchKernelGGL((k_residual<T, 3>), dim3(grid), tblock(), 0, s, la, oa, di[0], di[1],
                         di[2], rhs, u, res, part);
   else
      hipLaunchKernelGGL((k_residual<T, 2>), dim3(grid), tblock(), 0, s, la, oa, di[0], di[1],
                         di[2], rhs, u, res, part);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int launch_ell_resrestrict(const Level& coarse, const Level& fine, const EllOp<T>& op,
                           const double* dx, const T* rhs, const T* u, T* b, hipStream_t s)
{
   OpArgs<T> oa{op.D, op.Dc, op.C, op.Cc, op.M, op.Mc};
   T di[3];
   recip(dx, di);
   LevelArgs ca = level_args(coarse), fa = level_args(fine);
   if (fine.ndim == 3)
      hipLaunchKernelGGL((k_ell_resrestrict<T, 3>), dim3(coarse.ncell_tiles), tblock(), 0, s, ca,
                         fa, oa, di[0], di[1], di[2], rhs, u, b);
   else
      hipLaunchKernelGGL((k_ell_resrestrict<T, 2>), dim3(coarse.ncell_tiles), tblock(), 0, s, ca,
                         fa, oa, di[0], di[1], di[2], rhs, u, b);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int launch_quat_rbgs(const Level& L, const T* fc, const T* sm, double gamma, const double* dx,
                     int ncomp, const T* rhs, T* z, int color, hipStream_t s, bool split)
{
   LevelArgs la;
   int grid;
   shape(L, false, &la, &grid);
   T hi[3];
   recip2(dx, hi);
   KernelTimer::Span span(KernelTimer::QUAT_RBGS, s, quat_rbgs_bytes(L, ncomp, sizeof(T)));
   auto go = [&](auto ndim, auto ql) {
      constexpr int ND = decltype(ndim)::value, Q = decltype(ql)::value;
      if (split)
         hipLaunchKernelGGL((k_quat_rbgs_c<T, ND, Q, true>), dim3(grid), cblock(), 0, s, la, fc,
                            sm, (T)gamma, hi[0], hi[1], hi[2], ncomp, rhs, z, color);
      else
         hipLaunchKernelGGL((k_quat_rbgs_c<T, ND, Q, false>), dim3(grid), cblock(), 0, s, la, fc,
                            sm, (T)gamma, hi[0], hi[1], hi[2], ncomp, rhs, z, color);
   };
   quat_dispatch(L.ndim, ncomp, go);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int launch_quat_rbgs_fused(const Level& L, const T* fc, const T* sm, double gamma,
                           const double* dx, int ncomp, const T* rhs, const T* zin, T* zout,
                           hipStream_t s)
{
   if (ncomp > 4) {
      set_error("launch_quat_rbgs_fused: at most four components");
      return -1;
   }
   T hi[3];
   recip2(dx, hi);
   LevelArgs la = level_args(L);
   dim3 g(L.ncell_tiles);
   {
      KernelTimer::Span span(KernelTimer::QUAT_RBGS, s, quat_rbgs_fused_bytes(L, ncomp, sizeof(T)));
      LevelArgs lm = level_args_march(L);
      dim3 gm(L.nmarch_tiles);
      if (L.ndim == 3)
         hipLaunchKernelGGL((k_quat_rbgs_fused<T, 3>), gm, tblock(), 0, s, lm, fc, sm, (T)gamma,
                            hi[0], hi[1], hi[2], ncomp, rhs, zin, zout);
      else
         hipLaunchKernelGGL((k_quat_rbgs_fused<T, 2>), gm, tblock(), 0, s, lm, fc, sm, (T)gamma,
                            hi[0], hi[1], hi[2], ncomp, rhs, zin, zout);
      AMPE_HIP_CHECK(hipGetLastError());
   }
   if (ghost_fill(L, zout, ncomp, s)) return -1;
   if (L.ndim == 3)
      hipLaunchKernelGGL((k_quat_rbgs_shell<T, 3>), g, tblock(), 0, s, la, fc, sm, (T)gamma,
                         hi[0], hi[1], hi[2], ncomp, rhs, zin, zout);
   else
      hipLaunchKernelGGL((k_quat_rbgs_shell<T, 2>), g, tblock(), 0, s, la, fc, sm, (T)gamma,
                         hi[0], hi[1], hi[2], ncomp, rhs, zin, zout);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int launch_quat_residual(const Level& L, const T* fc, const T* sm, double gamma,
                         const double* dx, int ncomp, const T* rhs, const T* z, T* res,
                         double* part, hipStream_t s, bool split)
{
   LevelArgs la;
   int grid;
   shape(L, part != nullptr, &la, &grid);
   T hi[3];
   recip2(dx, hi);
   auto go = [&](auto ndim, auto ql) {
      constexpr int ND = decltype(ndim)::value, Q = decltype(ql)::value;
      if (split && !part)
         hipLaunchKernelGGL((k_quat_residual_p<T, ND, Q>), dim3(grid), cblock(), 0, s, la, fc,
                            sm, (T)gamma, hi[0], hi[1], hi[2], ncomp, rhs, z, res);
      else if (split)
         hipLaunchKernelGGL((k_quat_residual<T, ND, Q, true>), dim3(grid), tblock(), 0, s, la, fc,
                            sm, (T)gamma, hi[0], hi[1], hi[2], ncomp, rhs, z, res, part);
      else
         hipLaunchKernelGGL((k_quat_residual<T, ND, Q, false>), dim3(grid), tblock(), 0, s, la,
                            fc, sm, (T)gamma, hi[0], hi[1], hi[2], ncomp, rhs, z, res, part);
   };
   quat_dispatch(L.ndim, ncomp, go);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int launch_quat_resrestrict(const Level& coarse, const Level& fine, const T* fc, const T* sm,
                            double gamma, const double* dx, int ncomp, const T* rhs, const T* z,
                            T* b, hipStream_t s, bool split)
{
   if (ncomp != 1 && ncomp != 2 && ncomp != 4) {
      set_error("launch_quat_resrestrict: one, two or four components");
      return -1;
   }
   T hi[3];
   recip2(dx, hi);
   LevelArgs ca = level_args(coarse), fa = level_args(fine);
   auto go = [&](auto ndim, auto ql) {
      constexpr int ND = decltype(ndim)::value, Q = decltype(ql)::value;
      if constexpr (Q > 0) {
         if (split)
            hipLaunchKernelGGL((k_quat_resrestrict<T, ND, Q, true>), dim3(coarse.ncell_tiles),
                               tblock(), 0, s, ca, fa, fc, sm, (T)gamma, hi[0], hi[1], hi[2],
                               rhs, z, b);
         else
            hipLaunchKernelGGL((k_quat_resrestrict<T, ND, Q, false>), dim3(coarse.ncell_tiles),
                               tblock(), 0, s, ca, fa, fc, sm, (T)gamma, hi[0], hi[1], hi[2],
                               rhs, z, b);
      }
   };
   quat_dispatch(fine.ndim, ncomp, go);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

#define AMPE_INST(T)                                                                     \
   template int launch_rbgs<T>(const Level&, const EllOp<T>&, const double*, const T*, T*, \
                               int, bool, hipStream_t);                                  \
   template int launch_residual<T>(const Level&, const EllOp<T>&, const double*, const T*, \
                                   const T*, T*, double*, hipStream_t);                  \
   template int launch_quat_rbgs<T>(const Level&, const T*, const T*, double, const double*, \
                                    int, const T*, T*, int, hipStream_t, bool);          \
   template int launch_quat_residual<T>(const Level&, const T*, const T*, double,         \
                                        const double*, int, const T*, const T*, T*,      \
                                        double*, hipStream_t, bool);                     \
   template int launch_ell_resrestrict<T>(const Level&, const Level&, const EllOp<T>&,     \
                                          const double*, const T*, const T*, T*, hipStream_t); \
   template int launch_quat_resrestrict<T>(const Level&, const Level&, const T*, const T*, \
                                           double, const double*, int, const T*,         \
                                           const T*, T*, hipStream_t, bool);             \
   template int launch_quat_rbgs_fused<T>(const Level&, const T*, const T*, double,       \
                                          const double*, int, const T*, const T*, T*,    \
                                          hipStream_t);
AMPE_INST(double)
AMPE_INST(float)

}  // namespace ampe
