// Geometric multigrid below one patch level (see mg.h).
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "mg.h"

namespace ampe {

// ------------------------------------------------------------ transfers
// Coarse cell = mean of its 2^NDIM children (fixed summation order).
// SPLIT: f and c in the row parity-split layout (the two children of a row
// sit at the split positions of their columns; rows and planes as natural).
template <typename T, int NDIM, bool SPLIT>
__global__ void __launch_bounds__(256) k_restrict_cell(LevelArgs ca, LevelArgs fa, int ncomp,
                                                       const T* __restrict__ f,
                                                       T* __restrict__ c)
{
   CellPos p;
   if (!cell_pos(ca, 0, p)) return;
   const PatchDev fp = fa.P[ca.T[tile_index(ca)].patch];
   int fk = NDIM == 3 ? 2 * p.k : 0;
   const int ig = 2 * p.i + fa.g;
   const long long row = fp.cell_off + (long long)(2 * p.j + fa.g) * fp.sy +
                         (long long)(fk + fa.gz) * fp.sz;
   const long long f0 = SPLIT ? row + split_col(ig, fp.sy) : row + ig;
   const long long f1 = SPLIT ? row + split_col(ig + 1, fp.sy) : row + ig + 1;
   const long long co = SPLIT ? split_at(ca, p, 0) : p.c0;
   const long long sy = fp.sy, sz = fp.sz;
   for (int m = 0; m < ncomp; m++) {
      const T* fm = f + m * fa.ncg;
      T s = (fm[f0] + fm[f1]) + (fm[f0 + sy] + fm[f1 + sy]);
      if (NDIM == 3) s = s + ((fm[f0 + sz] + fm[f1 + sz]) + (fm[f0 + sz + sy] + fm[f1 + sz + sy]));
      c[m * ca.ncg + co] = s * (NDIM == 3 ? T(0.125) : T(0.25));
   }
}

// Coarse face = mean of the fine faces covering it.
template <typename T, int NDIM>
__global__ void __launch_bounds__(256) k_restrict_side(LevelArgs ca, LevelArgs fa,
                                                       const T* __restrict__ f,
                                                       T* __restrict__ c)
{
   CellPos p;
   if (!cell_pos(ca, 1, p)) return;
   if (NDIM == 2 && p.k > 0) return;
   const PatchDev fp = fa.P[ca.T[tile_index(ca)].patch];
   for (int d = 0; d < NDIM; d++) {
      bool ok = (p.i < p.pd.nx + (d == 0)) && (p.j < p.pd.ny + (d == 1)) &&
                (NDIM == 2 || p.k < p.pd.nz + (d == 2));
      if (!ok) continue;
      int b[3] = {2 * p.i, 2 * p.j, NDIM == 3 ? 2 * p.k : 0};
      int t1 = (d == 0) ? 1 : 0;  // first transverse direction
      int t2 = (d == 2) ? 1 : 2;  // second transverse direction (3-D)
      T s = 0;
      for (int a = 0; a < 2; a++)
         for (int e = 0; e < (NDIM == 3 ? 2 : 1); e++) {
            int q[3] = {b[0], b[1], b[2]};
            q[t1] += a;
            if (NDIM == 3) q[t2] += e;
            T v = f[face_idx(fa, fp, d, q[0], q[1], q[2])];
            s = (a == 0 && e == 0) ? v : s + v;
         }
      c[face_idx(ca, p.pd, d, p.i, p.j, p.k)] = s * (NDIM == 3 ? T(0.25) : T(0.5));
   }
}

// Fine cell += coarse correction: ORDER 1 tri/bilinear interpolation (weights
// 3/4 toward the parent, 1/4 toward the coarse neighbour on the cell's side in
// each direction; the coarse field needs filled ghosts), ORDER 0 the parent
// value (constant refinement).  SPLIT: xc and xf in the row parity-split
// layout.
template <typename T, int NDIM, int ORDER, bool SPLIT>
__device__ __forceinline__ void prolong_cell(const LevelArgs& fa, const LevelArgs& ca,
                                             const CellPos& p, int ncomp,
                                             const T* __restrict__ xc, T* __restrict__ xf)
{
   const PatchDev cp = ca.P[fa.T[tile_index(fa)].patch];
   int K = NDIM == 3 ? (p.k >> 1) : 0;
   const int ig = (p.i >> 1) + ca.g;
   const long long row = cp.cell_off + (long long)((p.j >> 1) + ca.g) * cp.sy +
                         (long long)(K + ca.gz) * cp.sz;
   const int di = (p.i & 1) ? 1 : -1;
   // the parent and its neighbour across the cell's side in the row
   const long long c0 = SPLIT ? row + split_col(ig, cp.sy) : row + ig;
   const long long c1 = SPLIT ? row + split_col(ig + di, cp.sy) : row + ig + di;
   long long oj = (p.j & 1) ? cp.sy : -cp.sy;
   long long ok = (NDIM == 3) ? ((p.k & 1) ? cp.sz : -cp.sz) : 0;
   const long long fo = SPLIT ? split_at(fa, p, 0) : p.c0;
   for (int m = 0; m < ncomp; m++) {
      const T* x = xc + m * ca.ncg;
      T v;
      if (ORDER == 0) {
         v = x[c0];
      } else if (NDIM == 2) {
         v = T(9.0 / 16.0) * x[c0] + T(3.0 / 16.0) * (x[c1] + x[c0 + oj]) +
             T(1.0 / 16.0) * x[c1 + oj];
      } else {
         v = T(27.0 / 64.0) * x[c0] +
             T(9.0 / 64.0) * ((x[c1] + x[c0 + oj]) + x[c0 + ok]) +
             T(3.0 / 64.0) * ((x[c1 + oj] + x[c1 + ok]) + x[c0 + oj + ok]) +
             T(1.0 / 64.0) * x[c1 + oj + ok];
      }
      T* y = xf + m * fa.ncg;
      y[fo] = y[fo] + v;
   }
}

template <typename T, int NDIM, int ORDER, bool SPLIT>
__global__ void __launch_bounds__(256) k_prolong_add(LevelArgs fa, LevelArgs ca, int ncomp,
                                                     const T* __restrict__ xc,
                                                     T* __restrict__ xf)
{
   CellPos p;
   if (cell_pos(fa, 0, p)) prolong_cell<T, NDIM, ORDER, SPLIT>(fa, ca, p, ncomp, xc, xf);
}

// The split-layout prolongation per coarse cell: a thread adds the
// correction to the 2^NDIM children of its coarse cell (consecutive lanes:
// consecutive coarse cells, so for each child offset the fine updates are
// contiguous in the split layout), each child's value by the formula and
// summation order of prolong_cell.
template <typename T, int NDIM, int ORDER>
__global__ void __launch_bounds__(256) k_prolong_add_cc(LevelArgs fa, LevelArgs ca, int ncomp,
                                                        const T* __restrict__ xc,
                                                        T* __restrict__ xf)
{
   CellPos p;
   if (!cell_pos(ca, 0, p)) return;
   const PatchDev fp = fa.P[ca.T[tile_index(ca)].patch];
   const int sy = p.pd.sy;
   const long long row = p.c0 - (p.i + ca.g);
   const int ig = p.i + ca.g;
   const long long c0 = row + split_col(ig, sy);
   const long long cm = row + split_col(ig - 1, sy), cp = row + split_col(ig + 1, sy);
   for (int m = 0; m < ncomp; m++) {
      const T* x = xc + m * ca.ncg;
      T* y = xf + m * fa.ncg;
#pragma unroll
      for (int dz = 0; dz < (NDIM == 3 ? 2 : 1); dz++)
#pragma unroll
         for (int dy = 0; dy < 2; dy++)
#pragma unroll
            for (int dx = 0; dx < 2; dx++) {
               const long long c1 = dx ? cp : cm;
               const long long oj = dy ? sy : -(long long)sy;
               const long long ok = NDIM == 3 ? (dz ? p.pd.sz : -p.pd.sz) : 0;
               T v;
               if (ORDER == 0) {
                  v = x[c0];
               } else if (NDIM == 2) {
                  v = T(9.0 / 16.0) * x[c0] + T(3.0 / 16.0) * (x[c1] + x[c0 + oj]) +
                      T(1.0 / 16.0) * x[c1 + oj];
               } else {
                  v = T(27.0 / 64.0) * x[c0] +
                      T(9.0 / 64.0) * ((x[c1] + x[c0 + oj]) + x[c0 + ok]) +
                      T(3.0 / 64.0) * ((x[c1 + oj] + x[c1 + ok]) + x[c0 + oj + ok]) +
                      T(1.0 / 64.0) * x[c1 + oj + ok];
               }
               const int fi = 2 * p.i + dx, fj = 2 * p.j + dy;
               const int fk = NDIM == 3 ? 2 * p.k + dz : 0;
               const long long fo = fp.cell_off + split_col(fi + fa.g, fp.sy) +
                                    (long long)(fj + fa.g) * fp.sy +
                                    (long long)(fk + fa.gz) * fp.sz;
               y[fo] = y[fo] + v;
            }
   }
}

template <typename T>
int mg_restrict_cell(const Level& coarse, const Level& fine, int ncomp, const T* f, T* c,
                     hipStream_t s, bool split)
{
   dim3 g(coarse.ncell_tiles), b(TX, TY);
   LevelArgs ca = level_args(coarse), fa = level_args(fine);
   auto go = [&](auto nd, auto sp) {
      hipLaunchKernelGGL((k_restrict_cell<T, decltype(nd)::value, decltype(sp)::value>), g, b, 0,
                         s, ca, fa, ncomp, f, c);
   };
   using I2 = std::integral_constant<int, 2>;
   using I3 = std::integral_constant<int, 3>;
   using ST = std::true_type;
   using SF = std::false_type;
   if (coarse.ndim == 3)
      split ? go(I3(), ST()) : go(I3(), SF());
   else
      split ? go(I2(), ST()) : go(I2(), SF());
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int mg_restrict_side(const Level& coarse, const Level& fine, const T* f, T* c, hipStream_t s)
{
   if (coarse.ndim == 3)
      hipLaunchKernelGGL((k_restrict_side<T, 3>), dim3(coarse.nface_tiles), dim3(TX, TY), 0, s,
                         level_args(coarse, true), level_args(fine), f, c);
   else
      hipLaunchKernelGGL((k_restrict_side<T, 2>), dim3(coarse.nface_tiles), dim3(TX, TY), 0, s,
                         level_args(coarse, true), level_args(fine), f, c);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int mg_prolong_add(const Level& fine, const Level& coarse, int ncomp, int order, const T* xc,
                   T* xf, hipStream_t s, bool split)
{
   dim3 g(fine.ncell_tiles), b(TX, TY);
   LevelArgs fa = level_args(fine), ca = level_args(coarse);
   auto go = [&](auto nd, auto sp) {
      constexpr int ND = decltype(nd)::value;
      constexpr bool SP = decltype(sp)::value;
      dim3 gc(coarse.ncell_tiles);
      if (SP && order == 0)
         hipLaunchKernelGGL((k_prolong_add_cc<T, ND, 0>), gc, b, 0, s, fa, ca, ncomp, xc, xf);
      else if (SP)
         hipLaunchKernelGGL((k_prolong_add_cc<T, ND, 1>), gc, b, 0, s, fa, ca, ncomp, xc, xf);
      else if (order == 0)
         hipLaunchKernelGGL((k_prolong_add<T, ND, 0, false>), g, b, 0, s, fa, ca, ncomp, xc, xf);
      else
         hipLaunchKernelGGL((k_prolong_add<T, ND, 1, false>), g, b, 0, s, fa, ca, ncomp, xc, xf);
   };
   using I2 = std::integral_constant<int, 2>;
   using I3 = std::integral_constant<int, 3>;
   using ST = std::true_type;
   using SF = std::false_type;
   if (fine.ndim == 3)
      split ? go(I3(), ST()) : go(I3(), SF());
   else
      split ? go(I2(), ST()) : go(I2(), SF());
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

bool is_coarsening(const Level& coarse, const Level& fine)
{
   if (coarse.ndim != fine.ndim || coarse.npatch != fine.npatch) return false;
   for (int p = 0; p < fine.npatch; p++) {
      const PatchDev& f = fine.hp[p];
      const PatchDev& c = coarse.hp[p];
      int nf[3] = {f.nx, f.ny, f.nz}, nc[3] = {c.nx, c.ny, c.nz};
      for (int a = 0; a < fine.ndim; a++)
         if ((f.lo[a] & 1) || (nf[a] & 1) || c.lo[a] * 2 != f.lo[a] || nc[a] * 2 != nf[a])
            return false;
   }
   return true;
}

// ------------------------------------------------------------ hierarchy
static bool halvable(const Level& L, int min_cells)
{
   int nd = L.ndim;
   for (const PatchDev& d : L.hp) {
      int n[3] = {d.nx, d.ny, d.nz};
      for (int a = 0; a < nd; a++) {
         if ((d.lo[a] & 1) || (n[a] & 1) || n[a] / 2 < min_cells) return false;
      }
   }
   if (L.periodic)
      for (int a = 0; a < nd; a++) {
         int ext = L.pbox.hi[a] - L.pbox.lo[a] + 1;
         if ((L.pbox.lo[a] & 1) || (ext & 1)) return false;
      }
   return true;
}

template <typename T>
static int talloc(T** p, size_t n)
{
   if (hipMalloc(p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return -1;
   return hipMemset(*p, 0, sizeof(T) * std::max<size_t>(n, 1)) == hipSuccess ? 0 : -1;
}

template <typename T>
int mg_create(const Level& fine, const double* dx, int ncomp_max, int max_levels, int min_cells,
              Mg<T>** out)
{
   Mg<T>* m = new Mg<T>();
   m->ncomp_max = ncomp_max;
   MgLevel<T> l0;
   l0.L = const_cast<Level*>(&fine);
   for (int a = 0; a < 3; a++) l0.dx[a] = dx[a];
   int rc = talloc(&l0.r, (size_t)ncomp_max * fine.ncg);
   rc |= talloc(&l0.t, (size_t)ncomp_max * fine.ncg);
   // ghost cells without a gather source are never written: keep them zero,
   // as in the arrays the sweeps alternate with
   if (!rc && hipMemset(l0.t, 0, sizeof(T) * (size_t)ncomp_max * fine.ncg) != hipSuccess) rc = -1;
   m->lv.push_back(l0);
   if (max_levels < 1) max_levels = 1;
   while (!rc && (int)m->lv.size() < max_levels && halvable(*m->lv.back().L, min_cells)) {
      const Level& F = *m->lv.back().L;
      std::vector<ampe_box> boxes(F.npatch);
      for (int p = 0; p < F.npatch; p++) {
         const PatchDev& d = F.hp[p];
         int n[3] = {d.nx, d.ny, d.nz};
         for (int a = 0; a < 3; a++) {
            boxes[p].lo[a] = a < F.ndim ? d.lo[a] / 2 : 0;
            boxes[p].hi[a] = a < F.ndim ? d.lo[a] / 2 + n[a] / 2 - 1 : 0;
         }
      }
      ampe_box pb;
      for (int a = 0; a < 3; a++) {
         pb.lo[a] = F.pbox.lo[a] / 2;
         pb.hi[a] = (F.pbox.hi[a] + 1) / 2 - 1;
      }
      MgLevel<T> c;
      if (level_create(F.ndim, boxes.data(), F.npatch, F.g, F.periodic ? &pb : nullptr, &c.L)) {
         rc = -1;
         break;
      }
      for (int a = 0; a < 3; a++) c.dx[a] = 2.0 * m->lv.back().dx[a];
      size_t ncg = (size_t)c.L->ncg, ns = (size_t)c.L->nside_tot;
      rc |= talloc(&c.C, ncg);
      rc |= talloc(&c.M, ncg);
      rc |= talloc(&c.D, ns);
      rc |= talloc(&c.fc, ns);
      rc |= talloc(&c.sm, ncg);
      rc |= talloc(&c.x, (size_t)ncomp_max * ncg);
      rc |= talloc(&c.b, (size_t)ncomp_max * ncg);
      rc |= talloc(&c.r, (size_t)ncomp_max * ncg);
      rc |= talloc(&c.t, (size_t)ncomp_max * ncg);
      if (!rc && hipMemset(c.t, 0, sizeof(T) * (size_t)ncomp_max * ncg) != hipSuccess) rc = -1;
      m->lv.push_back(c);
   }
   if (rc) {
      set_error("mg_create: out of device memory");
      mg_destroy(m);
      return -1;
   }
   *out = m;
   return 0;
}

template <typename T>
void mg_destroy(Mg<T>* m)
{
   if (!m) return;
   for (size_t l = 0; l < m->lv.size(); l++) {
      MgLevel<T>& c = m->lv[l];
      (void)hipFree(c.r);
      (void)hipFree(c.t);
      if (l == 0) continue;
      (void)hipFree(c.C);
      (void)hipFree(c.M);
      (void)hipFree(c.D);
      (void)hipFree(c.fc);
      (void)hipFree(c.sm);
      (void)hipFree(c.x);
      (void)hipFree(c.b);
      level_destroy(c.L);
   }
   delete m;
}

template <typename T>
int mg_setup_phase(Mg<T>& m, const T* C0, const T* M0, hipStream_t s)
{
   for (int l = 1; l < m.nlev(); l++) {
      const T* f = l == 1 ? C0 : m.lv[l - 1].C;
      if (mg_restrict_cell<T>(*m.lv[l].L, *m.lv[l - 1].L, 1, f, m.lv[l].C, s)) return -1;
      if (M0) {
         const T* fm = l == 1 ? M0 : m.lv[l - 1].M;
         if (mg_restrict_cell<T>(*m.lv[l].L, *m.lv[l - 1].L, 1, fm, m.lv[l].M, s)) return -1;
      }
   }
   return 0;
}

template <typename T>
int mg_setup_conc(Mg<T>& m, const T* D0, hipStream_t s)
{
   for (int l = 1; l < m.nlev(); l++) {
      const T* f = l == 1 ? D0 : m.lv[l - 1].D;
      if (mg_restrict_side<T>(*m.lv[l].L, *m.lv[l - 1].L, f, m.lv[l].D, s)) return -1;
   }
   return 0;
}

template <typename T>
int mg_setup_quat(Mg<T>& m, const T* fc0, const T* sm0, hipStream_t s)
{
   for (int l = 1; l < m.nlev(); l++) {
      const Level& C = *m.lv[l].L;
      const Level& F = *m.lv[l - 1].L;
      const T* ff = l == 1 ? fc0 : m.lv[l - 1].fc;
      const T* fs = l == 1 ? sm0 : m.lv[l - 1].sm;
      if (mg_restrict_side<T>(C, F, ff, m.lv[l].fc, s)) return -1;
      if (mg_restrict_cell<T>(C, F, 1, fs, m.lv[l].sm, s)) return -1;
      if (ghost_fill(C, m.lv[l].sm, 1, s)) return -1;
   }
   return 0;
}

// ------------------------------------------------------------ cycles
template <typename T>
static int smooth_ell(const Level& L, const EllOp<T>& op, const double* dx, const T* b, T* x,
                      int sweeps, hipStream_t s)
{
   for (int sw = 0; sw < sweeps; sw++) {
      if (ghost_fill(L, x, 1, s)) return -1;
      if (launch_rbgs<T>(L, op, dx, b, x, 0, false, s)) return -1;
      if (ghost_fill(L, x, 1, s)) return -1;
      if (launch_rbgs<T>(L, op, dx, b, x, 1, false, s)) return -1;
   }
   return 0;
}

template <typename T>
static int smooth_quat(const Level& L, const T* fc, const T* sm, double gamma, const double* dx,
                       int ncomp, const T* b, T* x, int sweeps, hipStream_t s, bool split)
{
   for (int sw = 0; sw < sweeps; sw++) {
      if (ghost_fill(L, x, ncomp, s, split)) return -1;
      if (launch_quat_rbgs<T>(L, fc, sm, gamma, dx, ncomp, b, x, 0, s, split)) return -1;
      if (ghost_fill(L, x, ncomp, s, split)) return -1;
      if (launch_quat_rbgs<T>(L, fc, sm, gamma, dx, ncomp, b, x, 1, s, split)) return -1;
   }
   return 0;
}

// nsweeps fused sweeps alternating between *cur and *oth; on return *cur
// holds the result.
template <typename T>
static int smooth_quat_pp(const Level& L, const T* fc, const T* sm, double gamma, const double* dx,
                          int ncomp, const T* b, T** cur, T** oth, int sweeps, hipStream_t s)
{
   for (int sw = 0; sw < sweeps; sw++) {
      if (ghost_fill(L, *cur, ncomp, s)) return -1;
      if (launch_quat_rbgs_fused<T>(L, fc, sm, gamma, dx, ncomp, b, *cur, *oth, s)) return -1;
      std::swap(*cur, *oth);
   }
   return 0;
}

template <typename T>
static EllOp<T> level_op(const Mg<T>& m, const EllOp<T>& op0, int kind, int l)
{
   EllOp<T> op = op0;
   if (l > 0) {
      if (kind == 0) {
         op.C = m.lv[l].C;
         if (op0.M) op.M = m.lv[l].M;
      } else {
         op.D = m.lv[l].D;
      }
   }
   return op;
}

template <typename T>
int mg_vcycle_ell(Mg<T>& m, const EllOp<T>& op0, int kind, const MgCycle& cy, const T* b0,
                  T* x0, hipStream_t s)
{
   if (op0.M && kind != 0) {
      set_error("mg_vcycle_ell: a mobility field is supported for the phase block only");
      return -1;
   }
   int nl = m.nlev();
   // descend
   for (int l = 0; l < nl; l++) {
      MgLevel<T>& c = m.lv[l];
      const T* b = l == 0 ? b0 : c.b;
      T* x = l == 0 ? x0 : c.x;
      EllOp<T> op = level_op(m, op0, kind, l);
      if (l == nl - 1) {
         if (smooth_ell<T>(*c.L, op, c.dx, b, x, cy.ncoarse, s)) return -1;
         break;
      }
      if (smooth_ell<T>(*c.L, op, c.dx, b, x, cy.nu1, s)) return -1;
      if (ghost_fill(*c.L, x, 1, s)) return -1;
      MgLevel<T>& n = m.lv[l + 1];
      if (cy.nu1 >= 1 && red_residual_ok(*c.L, 1)) {
         // the sweeps end with the black colour: restrict the red residuals
         if (launch_ell_resrestrict<T>(*n.L, *c.L, op, c.dx, b, x, n.b, s)) return -1;
      } else {
         if (launch_residual<T>(*c.L, op, c.dx, b, x, c.r, nullptr, s)) return -1;
         if (mg_restrict_cell<T>(*n.L, *c.L, 1, c.r, n.b, s)) return -1;
      }
      if (fill_zero<T>(n.x, n.L->ncg, s)) return -1;
   }
   // ascend
   for (int l = nl - 2; l >= 0; l--) {
      MgLevel<T>& c = m.lv[l];
      MgLevel<T>& n = m.lv[l + 1];
      const T* b = l == 0 ? b0 : c.b;
      T* x = l == 0 ? x0 : c.x;
      if (ghost_fill(*n.L, n.x, 1, s)) return -1;
      if (mg_prolong_add<T>(*c.L, *n.L, 1, 1, n.x, x, s)) return -1;
      if (smooth_ell<T>(*c.L, level_op(m, op0, kind, l), c.dx, b, x, cy.nu2, s)) return -1;
   }
   return 0;
}

// The residual after red-black sweeps ending with the black colour vanishes
// on black cells when every black cell's neighbours are red: one checkerboard
// over the patches (non-negative, even patch corners: the colour rule of the
// sweeps then depends on global indices only) and periodic ghosts (zero-slope
// ghosts copy a cell's own value into its row).  One, two or four components.
bool red_residual_ok(const Level& L, int ncomp)
{
   if (!L.periodic || (ncomp != 1 && ncomp != 2 && ncomp != 4)) return false;
   for (int a = 0; a < L.ndim; a++) {
      int ext = L.pbox.hi[a] - L.pbox.lo[a] + 1;
      if (L.pbox.lo[a] < 0 || (L.pbox.lo[a] & 1) || (ext & 1)) return false;
   }
   for (const PatchDev& d : L.hp)
      for (int a = 0; a < L.ndim; a++)
         if (d.lo[a] < 0 || (d.lo[a] & 1)) return false;
   return true;
}

template <typename T>
int mg_vcycle_quat(Mg<T>& m, const T* fc0, const T* sm0, double gamma, int ncomp,
                   const MgCycle& cy, const T* b0, T* x0, hipStream_t s)
{
   if (ncomp > m.ncomp_max) {
      set_error("mg_vcycle_quat: more components than the hierarchy holds");
      return -1;
   }
   int nl = m.nlev();
   if (cy.fused && m.qsplit) {
      set_error("mg_vcycle_quat: the one-pass sweeps use the natural layout");
      return -1;
   }
   if (cy.fused) {
      // one-pass sweeps alternate between x and t on every level; cur[l] is
      // the level's current iterate
      std::vector<T*> cur(nl), oth(nl);
      for (int l = 0; l < nl; l++) {
         cur[l] = l == 0 ? x0 : m.lv[l].x;
         oth[l] = m.lv[l].t;
      }
      for (int l = 0; l < nl; l++) {
         MgLevel<T>& c = m.lv[l];
         const T* b = l == 0 ? b0 : c.b;
         const T* fc = l == 0 ? fc0 : c.fc;
         const T* sm = l == 0 ? sm0 : c.sm;
         int nsw = (l == nl - 1) ? cy.ncoarse : cy.nu1;
         if (smooth_quat_pp<T>(*c.L, fc, sm, gamma, c.dx, ncomp, b, &cur[l], &oth[l], nsw, s))
            return -1;
         if (l == nl - 1) break;
         if (ghost_fill(*c.L, cur[l], ncomp, s)) return -1;
         if (launch_quat_residual<T>(*c.L, fc, sm, gamma, c.dx, ncomp, b, cur[l], c.r, nullptr,
                                     s))
            return -1;
         MgLevel<T>& n = m.lv[l + 1];
         if (mg_restrict_cell<T>(*n.L, *c.L, ncomp, c.r, n.b, s)) return -1;
         if (fill_zero<T>(cur[l + 1], n.L->ncg * ncomp, s)) return -1;
      }
      for (int l = nl - 2; l >= 0; l--) {
         MgLevel<T>& c = m.lv[l];
         MgLevel<T>& n = m.lv[l + 1];
         const T* b = l == 0 ? b0 : c.b;
         const T* fc = l == 0 ? fc0 : c.fc;
         const T* sm = l == 0 ? sm0 : c.sm;
         if (ghost_fill(*n.L, cur[l + 1], ncomp, s)) return -1;
         if (mg_prolong_add<T>(*c.L, *n.L, ncomp, 1, cur[l + 1], cur[l], s)) return -1;
         if (smooth_quat_pp<T>(*c.L, fc, sm, gamma, c.dx, ncomp, b, &cur[l], &oth[l], cy.nu2, s))
            return -1;
      }
      if (cur[0] != x0)
         AMPE_HIP_CHECK(hipMemcpyAsync(x0, cur[0], sizeof(T) * m.lv[0].L->ncg * ncomp,
                                       hipMemcpyDeviceToDevice, s));
      return 0;
   }
   const bool sp = m.qsplit;
   for (int l = 0; l < nl; l++) {
      MgLevel<T>& c = m.lv[l];
      const T* b = l == 0 ? b0 : c.b;
      T* x = l == 0 ? x0 : c.x;
      const T* fc = l == 0 ? fc0 : c.fc;
      const T* sm = l == 0 ? sm0 : c.sm;
      if (l == nl - 1) {
         if (smooth_quat<T>(*c.L, fc, sm, gamma, c.dx, ncomp, b, x, cy.ncoarse, s, sp)) return -1;
         break;
      }
      if (smooth_quat<T>(*c.L, fc, sm, gamma, c.dx, ncomp, b, x, cy.nu1, s, sp)) return -1;
      if (ghost_fill(*c.L, x, ncomp, s, sp)) return -1;
      MgLevel<T>& n = m.lv[l + 1];
      if (cy.nu1 >= 1 && red_residual_ok(*c.L, ncomp)) {
         // the sweeps end with the black colour: restrict the red residuals
         if (launch_quat_resrestrict<T>(*n.L, *c.L, fc, sm, gamma, c.dx, ncomp, b, x, n.b, s, sp))
            return -1;
      } else {
         if (launch_quat_residual<T>(*c.L, fc, sm, gamma, c.dx, ncomp, b, x, c.r, nullptr, s,
                                     sp))
            return -1;
         if (mg_restrict_cell<T>(*n.L, *c.L, ncomp, c.r, n.b, s, sp)) return -1;
      }
      if (fill_zero<T>(n.x, n.L->ncg * ncomp, s)) return -1;
   }
   for (int l = nl - 2; l >= 0; l--) {
      MgLevel<T>& c = m.lv[l];
      MgLevel<T>& n = m.lv[l + 1];
      const T* b = l == 0 ? b0 : c.b;
      T* x = l == 0 ? x0 : c.x;
      const T* fc = l == 0 ? fc0 : c.fc;
      const T* sm = l == 0 ? sm0 : c.sm;
      if (ghost_fill(*n.L, n.x, ncomp, s, sp)) return -1;
      if (mg_prolong_add<T>(*c.L, *n.L, ncomp, 1, n.x, x, s, sp)) return -1;
      if (smooth_quat<T>(*c.L, fc, sm, gamma, c.dx, ncomp, b, x, cy.nu2, s, sp)) return -1;
   }
   return 0;
}

template <typename T>
int mg_enable_qsplit(Mg<T>& m, hipStream_t s)
{
   for (auto& c : m.lv)
      if (level_split_tables(*c.L, s)) return -1;
   m.qsplit = true;
   return 0;
}

#define AMPE_MG_INST(T)                                                                      \
   template int mg_create<T>(const Level&, const double*, int, int, int, Mg<T>**);          \
   template void mg_destroy<T>(Mg<T>*);                                                      \
   template int mg_setup_phase<T>(Mg<T>&, const T*, const T*, hipStream_t);                  \
   template int mg_setup_conc<T>(Mg<T>&, const T*, hipStream_t);                             \
   template int mg_setup_quat<T>(Mg<T>&, const T*, const T*, hipStream_t);                   \
   template int mg_vcycle_ell<T>(Mg<T>&, const EllOp<T>&, int, const MgCycle&, const T*, T*, \
                                 hipStream_t);                                               \
   template int mg_vcycle_quat<T>(Mg<T>&, const T*, const T*, double, int, const MgCycle&,   \
                                  const T*, T*, hipStream_t);                                \
   template int mg_restrict_cell<T>(const Level&, const Level&, int, const T*, T*,           \
                                    hipStream_t, bool);                                      \
   template int mg_enable_qsplit<T>(Mg<T>&, hipStream_t);                                    \
   template int mg_restrict_side<T>(const Level&, const Level&, const T*, T*, hipStream_t);  \
   template int mg_prolong_add<T>(const Level&, const Level&, int, int, const T*, T*,        \
                                  hipStream_t, bool);
AMPE_MG_INST(float)
AMPE_MG_INST(double)

}  // namespace ampe
