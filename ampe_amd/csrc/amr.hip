// A hierarchy of patch levels (refinement ratio 2) on the device.
//
// Level 0 covers the domain; level l+1 is a set of boxes in its own index
// space, aligned with level-l cells and properly nested.  Every field of the
// hierarchy is one composite array: per component, the ghosted blocks of all
// patches of all levels, level by level; side fields hold, per direction, the
// faces of all patches.  Each level is a Level placed in those arrays, so
// every per-level kernel of the library runs on it unchanged, and a
// composite Level over all patches carries the whole-vector operations.
//
// The transfers between levels run from tables built once on the host:
//  - coarse-fine ghosts: the coarse parent (CONSTANT_REFINE) or the four /
//    eight coarse cells and weights of quatlinrefine's interpolation
//    (LINEAR_REFINE), the refine schedules of QuatIntegrator::fillScratch
//    (source/QuatIntegrator.cc:2873-2957, operators QuatModel.cc:924-955);
//  - CONSERVATIVE_COARSEN of cells (QuatIntegrator::coarsenData, :2289-2331)
//    and of side data (the flux and diffusion-coefficient coarsen schedules,
//    :945-955, :2424-2427);
//  - the coarse-fine faces of each fine level and, on the next coarser
//    level, the uncovered cells next to them, whose divergence takes the
//    coarsened fine fluxes (composite right-hand sides and the composite
//    residual of the FAC preconditioner);
//  - the FAC cycle over levels (FACPreconditioner::solveSystem /
//    facCycle_Recursive, source/samrai/FACPreconditioner.cc:267-474) with
//    EllipticFACOps' smoothing, residual, restriction, constant-refine
//    prolongation and Ewing's coarse-fine flux fix (source/
//    EllipticFACOps.cc:948-1241, 1480-1760), the fix applied through an
//    effective ghost value per coarse-fine face.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "amr.h"
#include "fac.h"
#include "kernels.h"

namespace ampe {

constexpr int kMaxLev = 4;

template <typename T>
struct DevVec {
   T* p = nullptr;
   long long n = 0;
   int upload(const std::vector<T>& h)
   {
      n = (long long)h.size();
      if (n == 0) return 0;
      AMPE_HIP_CHECK(hipMalloc((void**)&p, sizeof(T) * h.size()));
      AMPE_HIP_CHECK(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
      return 0;
   }
   void release()
   {
      if (p) (void)hipFree(p);
      p = nullptr;
      n = 0;
   }
};

struct Hier {
   int ndim = 2, nlev = 1, periodic = 0;
   int dom[kMaxLev][3];
   double dx[kMaxLev][3];
   int pbeg[kMaxLev + 1];
   std::vector<ampe_box> boxes;
   Level* comp = nullptr;
   Level* lev[kMaxLev] = {};
   long long ncg = 0, ncell = 0, nside[3] = {0, 0, 0}, nside_tot = 0;
   std::vector<long long> coff, int_off;  // per patch (host)
   std::vector<long long> soff[3];       // per patch face offset in direction block
   // coarse-fine ghosts of level l (l > 0): dst, 9 sources, 3 weights
   DevVec<long long> cf_dst[kMaxLev], cf_src[kMaxLev];
   DevVec<double> cf_w[kMaxLev];
   // cell coarsening onto level l (l < nlev-1): coarse cell; its first fine
   // cell and the fine patch's row and plane strides (low / high 32 bits)
   DevVec<long long> crs_dst[kMaxLev], crs_src[kMaxLev];
   // face coarsening onto level l: coarse face, 4 fine faces, direction
   DevVec<long long> fcrs_dst[kMaxLev], fcrs_src[kMaxLev];
   DevVec<int> fcrs_dir[kMaxLev];
   // coarse-fine faces of level l (l > 0) and corrections of level l (< nlev-1)
   DevVec<CfFace> cff[kMaxLev];
   DevVec<CfCorr> cfc[kMaxLev];
   // the FAC composite residual replaces every coarse face on a fine patch
   // boundary: fine faces of level l on fine-fine patch boundaries (saved
   // after cff[l]) and the covered coarse cells of level l next to a fine
   // patch boundary of level l+1
   DevVec<CfFace> fff[kMaxLev];
   DevVec<CfCorr> cfi[kMaxLev];
   // prolongation of level l (l > 0): fine interior cell, coarse parent
   DevVec<long long> pro_dst[kMaxLev], pro_src[kMaxLev];
   double* wt = nullptr;     // composite weights (interior cells), device
   // FAC scratch (grown on demand)
   FacWork<double> wd;  // FAC work arrays of the FP64 block solves
   FacWork<float> wf;   // and of the FP32 ones
};

namespace {

int fdiv2(int a) { return a < 0 ? (a + 1) / 2 - 1 : a / 2; }

void canon(const Hier& H, int l, int g[3])
{
   for (int a = 0; a < H.ndim; a++) {
      int n = H.dom[l][a];
      if (H.periodic)
         g[a] = ((g[a] % n) + n) % n;
      else
         g[a] = std::min(std::max(g[a], 0), n - 1);
   }
   if (H.ndim == 2) g[2] = 0;
}

int box_n(const ampe_box& b, int a, int ndim) { return a < ndim ? b.hi[a] - b.lo[a] + 1 : 1; }

bool inbox(const Hier& H, int p, const int g[3])
{
   const ampe_box& b = H.boxes[p];
   for (int a = 0; a < 3; a++) {
      int lo = a < H.ndim ? b.lo[a] : 0;
      if (g[a] < lo || g[a] >= lo + box_n(b, a, H.ndim)) return false;
   }
   return true;
}

// ghosted-cell composite index of patch p's local cell (i, j, k)
long long cidx(const Hier& H, int p, int i, int j, int k)
{
   const ampe_box& b = H.boxes[p];
   long long sy = box_n(b, 0, H.ndim) + 2;
   long long sz = sy * (box_n(b, 1, H.ndim) + 2);
   int gz = H.ndim == 3 ? 1 : 0;
   return H.coff[p] + (i + 1) + (long long)(j + 1) * sy + (long long)(k + gz) * sz;
}

long long find_cell(const Hier& H, int l, const int g[3], int* pp = nullptr)
{
   for (int p = H.pbeg[l]; p < H.pbeg[l + 1]; p++)
      if (inbox(H, p, g)) {
         if (pp) *pp = p;
         const ampe_box& b = H.boxes[p];
         return cidx(H, p, g[0] - b.lo[0], g[1] - b.lo[1], H.ndim == 3 ? g[2] - b.lo[2] : 0);
      }
   return -1;
}

// composite side index of patch p's face (i, j, k) of direction d
long long fidx(const Hier& H, int p, int d, int i, int j, int k)
{
   const ampe_box& b = H.boxes[p];
   int ex = box_n(b, 0, H.ndim) + (d == 0), ey = box_n(b, 1, H.ndim) + (d == 1);
   long long dir_off = d == 0 ? 0 : (d == 1 ? H.nside[0] : H.nside[0] + H.nside[1]);
   return dir_off + H.soff[d][p] + i + (long long)ex * (j + (long long)ey * k);
}

void lin_index(int gf, double dxf, double dxc, int& gc, double& x)
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

}  // namespace

void hier_destroy(Hier* H)
{
   if (!H) return;
   for (int l = 0; l < kMaxLev; l++) {
      if (H->lev[l]) level_destroy(H->lev[l]);
      H->cf_dst[l].release();
      H->cf_src[l].release();
      H->cf_w[l].release();
      H->crs_dst[l].release();
      H->crs_src[l].release();
      H->fcrs_dst[l].release();
      H->fcrs_src[l].release();
      H->fcrs_dir[l].release();
      H->cff[l].release();
      H->cfc[l].release();
      H->fff[l].release();
      H->cfi[l].release();
      H->pro_dst[l].release();
      H->pro_src[l].release();
   }
   if (H->comp) level_destroy(H->comp);
   (void)hipFree(H->wt);
   H->wd.release();
   H->wf.release();
   delete H;
}

static int hier_fail(Hier* H, const char* msg)
{
   set_error(std::string("hier_create: ") + msg);
   hier_destroy(H);
   return -1;
}

int hier_create(int ndim, int nlev, const int* npatch, const ampe_box* boxes, const int dom0[3],
                const double dx0[3], int periodic, Hier** out)
{
   if ((ndim != 2 && ndim != 3) || nlev < 1 || nlev > kMaxLev || !npatch || !boxes || !dom0 ||
       !dx0) {
      set_error("hier_create: bad arguments");
      return -1;
   }
   for (int a = 1; a < ndim; a++)
      if (dx0[a] != dx0[0]) {
         set_error("hier_create: the mesh spacing must be the same in every direction");
         return -1;
      }
   Hier* H = new Hier();
   H->ndim = ndim;
   H->nlev = nlev;
   H->periodic = periodic ? 1 : 0;
   int np = 0;
   for (int l = 0; l < nlev; l++) {
      if (npatch[l] <= 0) return hier_fail(H, "every level needs a patch");
      H->pbeg[l] = np;
      np += npatch[l];
      for (int a = 0; a < 3; a++) {
         H->dom[l][a] = a < ndim ? dom0[a] << l : 1;
         H->dx[l][a] = a < ndim ? dx0[a] / (double)(1 << l) : 1.0;
      }
   }
   H->pbeg[nlev] = np;
   H->boxes.assign(boxes, boxes + np);
   for (int p = 0; p < np; p++)
      if (ndim == 2) H->boxes[p].lo[2] = H->boxes[p].hi[2] = 0;
   // validation: inside the domain, aligned above level 0, disjoint per
   // level, level 0 covering the domain
   for (int l = 0; l < nlev; l++) {
      long long cells = 0;
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         const ampe_box& b = H->boxes[p];
         for (int a = 0; a < ndim; a++) {
            int n = box_n(b, a, ndim);
            if (n < 1 || b.lo[a] < 0 || b.lo[a] + n > H->dom[l][a])
               return hier_fail(H, "box outside its level's domain");
            if (l > 0 && ((b.lo[a] & 1) || (n & 1)))
               return hier_fail(H, "fine boxes must be aligned with the coarser cells");
         }
         cells += (long long)box_n(b, 0, ndim) * box_n(b, 1, ndim) * box_n(b, 2, ndim);
         for (int q = H->pbeg[l]; q < p; q++) {
            const ampe_box& c = H->boxes[q];
            bool ov = true;
            for (int a = 0; a < ndim; a++)
               if (b.hi[a] < c.lo[a] || c.hi[a] < b.lo[a]) ov = false;
            if (ov) return hier_fail(H, "boxes of one level overlap");
         }
      }
      if (l == 0 && cells != (long long)H->dom[0][0] * H->dom[0][1] * H->dom[0][2])
         return hier_fail(H, "level 0 must cover the domain");
   }
   // composite layout
   H->coff.resize(np);
   H->int_off.resize(np);
   for (int a = 0; a < 3; a++) H->soff[a].resize(np);
   long long co = 0, io = 0, so[3] = {0, 0, 0};
   for (int p = 0; p < np; p++) {
      const ampe_box& b = H->boxes[p];
      int n[3] = {box_n(b, 0, ndim), box_n(b, 1, ndim), box_n(b, 2, ndim)};
      H->coff[p] = co;
      H->int_off[p] = io;
      co += (long long)(n[0] + 2) * (n[1] + 2) * (n[2] + (ndim == 3 ? 2 : 0));
      io += (long long)n[0] * n[1] * n[2];
      for (int a = 0; a < 3; a++) {
         H->soff[a][p] = so[a];
         if (a < ndim) so[a] += (long long)(n[0] + (a == 0)) * (n[1] + (a == 1)) * (n[2] + (a == 2));
      }
   }
   H->ncg = co;
   H->ncell = io;
   for (int a = 0; a < 3; a++) H->nside[a] = so[a];
   H->nside_tot = so[0] + so[1] + so[2];

   // ghost tables per level
   std::vector<long long> gd[kMaxLev], gs[kMaxLev], cd[kMaxLev], cs9[kMaxLev];
   std::vector<double> cw[kMaxLev];
   for (int l = 0; l < nlev; l++)
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         const ampe_box& b = H->boxes[p];
         int n[3] = {box_n(b, 0, ndim), box_n(b, 1, ndim), box_n(b, 2, ndim)};
         int gz = ndim == 3 ? 1 : 0;
         for (int k = -gz; k < n[2] + gz; k++)
            for (int j = -1; j < n[1] + 1; j++)
               for (int i = -1; i < n[0] + 1; i++) {
                  if (i >= 0 && i < n[0] && j >= 0 && j < n[1] && k >= 0 && k < n[2]) continue;
                  int g[3] = {b.lo[0] + i, b.lo[1] + j, ndim == 3 ? b.lo[2] + k : 0};
                  canon(*H, l, g);
                  long long dst = cidx(*H, p, i, j, k);
                  long long s = find_cell(*H, l, g);
                  if (s >= 0) {
                     gd[l].push_back(dst);
                     gs[l].push_back(s);
                     continue;
                  }
                  if (l == 0) return hier_fail(H, "level 0 leaves a ghost uncovered");
                  int cp[3] = {fdiv2(g[0]), fdiv2(g[1]), ndim == 3 ? fdiv2(g[2]) : 0};
                  canon(*H, l - 1, cp);
                  long long sc = find_cell(*H, l - 1, cp);
                  if (sc < 0) return hier_fail(H, "a level is not nested in the coarser one");
                  cd[l].push_back(dst);
                  cs9[l].push_back(sc);
                  int gc[3] = {0, 0, 0};
                  double w[3] = {0.0, 0.0, 0.0};
                  for (int a = 0; a < ndim; a++)
                     lin_index(g[a], H->dx[l][a], H->dx[l - 1][a], gc[a], w[a]);
                  for (int a = 0; a < 3; a++) cw[l].push_back(w[a]);
                  static const int off[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0},
                                                {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
                  for (int c = 0; c < 8; c++) {
                     if (c >= (ndim == 3 ? 8 : 4)) {
                        cs9[l].push_back(sc);
                        continue;
                     }
                     int cc[3] = {gc[0] + off[c][0], gc[1] + off[c][1],
                                  ndim == 3 ? gc[2] + off[c][2] : 0};
                     canon(*H, l - 1, cc);
                     long long sv = find_cell(*H, l - 1, cc);
                     if (sv < 0) return hier_fail(H, "a level is not nested in the coarser one");
                     cs9[l].push_back(sv);
                  }
               }
      }

   // levels placed in the composite layout, and the composite over all patches
   for (int l = 0; l < nlev; l++) {
      LevelPlacement pl;
      int p0 = H->pbeg[l];
      pl.cell_base = H->coff[p0];
      pl.ncg_total = H->ncg;
      pl.int_base = H->int_off[p0];
      pl.ncell_total = H->ncell;
      for (int a = 0; a < 3; a++) {
         pl.side_base[a] = H->soff[a][p0];
         pl.nside_total[a] = H->nside[a];
      }
      if (level_create_placed(ndim, H->boxes.data() + p0, H->pbeg[l + 1] - p0, pl, gd[l], gs[l],
                              &H->lev[l]))
         return hier_fail(H, "level placement failed");
   }
   {
      LevelPlacement pl;
      pl.ncg_total = H->ncg;
      pl.ncell_total = H->ncell;
      for (int a = 0; a < 3; a++) pl.nside_total[a] = H->nside[a];
      std::vector<long long> none;
      if (level_create_placed(ndim, H->boxes.data(), np, pl, none, none, &H->comp))
         return hier_fail(H, "composite placement failed");
   }
   for (int l = 1; l < nlev; l++) {
      if (H->cf_dst[l].upload(cd[l]) || H->cf_src[l].upload(cs9[l]) || H->cf_w[l].upload(cw[l]))
         return hier_fail(H, "upload failed");
   }

   // cell and face coarsening, prolongation, coarse-fine faces, corrections
   for (int l = 0; l + 1 < nlev; l++) {
      std::vector<long long> dst, src, fdst, fsrc, pdst, psrc;
      std::vector<int> fdir;
      for (int p = H->pbeg[l + 1]; p < H->pbeg[l + 2]; p++) {
         const ampe_box& b = H->boxes[p];
         int cn[3] = {box_n(b, 0, ndim) / 2, box_n(b, 1, ndim) / 2,
                      ndim == 3 ? box_n(b, 2, ndim) / 2 : 1};
         int clo[3] = {b.lo[0] / 2, b.lo[1] / 2, ndim == 3 ? b.lo[2] / 2 : 0};
         for (int ck = 0; ck < cn[2]; ck++)
            for (int cj = 0; cj < cn[1]; cj++)
               for (int ci = 0; ci < cn[0]; ci++) {
                  int g[3] = {clo[0] + ci, clo[1] + cj, clo[2] + ck};
                  long long c0 = find_cell(*H, l, g);
                  if (c0 < 0) return hier_fail(H, "a level is not nested in the coarser one");
                  dst.push_back(c0);
                  // the first child and the fine patch's strides: child
                  // (r0, r1, r2) is base + r0 + r1 sy + r2 sz
                  const long long base = cidx(*H, p, 2 * ci, 2 * cj, ndim == 3 ? 2 * ck : 0);
                  const long long sy = cidx(*H, p, 2 * ci, 2 * cj + 1, ndim == 3 ? 2 * ck : 0) - base;
                  const long long sz =
                      ndim == 3 ? cidx(*H, p, 2 * ci, 2 * cj, 2 * ck + 1) - base : 0;
                  src.push_back(base);
                  src.push_back(sy | (sz << 32));
               }
         // faces of the coarsened box, written into every coarse patch holding them
         for (int d = 0; d < ndim; d++) {
            int t1 = (ndim == 2) ? 1 - d : (d + 1) % 3, t2 = (ndim == 2) ? 2 : (d + 2) % 3;
            int e[3] = {cn[0] + (d == 0), cn[1] + (d == 1), cn[2] + (d == 2)};
            for (int ck = 0; ck < e[2]; ck++)
               for (int cj = 0; cj < e[1]; cj++)
                  for (int ci = 0; ci < e[0]; ci++) {
                     int cf[3] = {ci, cj, ck};
                     int gcf[3] = {clo[0] + ci, clo[1] + cj, clo[2] + ck};
                     long long fine[4];
                     int nr = 0;
                     for (int r2 = 0; r2 < (ndim == 3 ? 2 : 1); r2++)
                        for (int r1 = 0; r1 < 2; r1++) {
                           int ff[3] = {0, 0, 0};
                           ff[d] = 2 * cf[d];
                           ff[t1] = 2 * cf[t1] + r1;
                           if (ndim == 3) ff[t2] = 2 * cf[t2] + r2;
                           fine[nr++] = fidx(*H, p, d, ff[0], ff[1], ff[2]);
                        }
                     for (int q = H->pbeg[l]; q < H->pbeg[l + 1]; q++) {
                        const ampe_box& c = H->boxes[q];
                        int loc[3];
                        bool ok = true;
                        for (int a = 0; a < 3; a++) {
                           int lo = a < ndim ? c.lo[a] : 0;
                           loc[a] = gcf[a] - lo;
                           int ext = box_n(c, a, ndim) + (a == d);
                           if (loc[a] < 0 || loc[a] >= ext) ok = false;
                        }
                        if (!ok) continue;
                        fdst.push_back(fidx(*H, q, d, loc[0], loc[1], loc[2]));
                        for (int r = 0; r < 4; r++) fsrc.push_back(r < nr ? fine[r] : -1);
                        fdir.push_back(d);
                     }
                  }
         }
      }
      if (H->crs_dst[l].upload(dst) || H->crs_src[l].upload(src) ||
          H->fcrs_dst[l].upload(fdst) || H->fcrs_src[l].upload(fsrc) ||
          H->fcrs_dir[l].upload(fdir))
         return hier_fail(H, "upload failed");
   }
   for (int l = 1; l < nlev; l++) {
      std::vector<long long> pdst, psrc;
      std::vector<CfFace> faces;
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         const ampe_box& b = H->boxes[p];
         int n[3] = {box_n(b, 0, ndim), box_n(b, 1, ndim), box_n(b, 2, ndim)};
         long long sy = n[0] + 2, sz = sy * (n[1] + 2);
         for (int k = 0; k < n[2]; k++)
            for (int j = 0; j < n[1]; j++)
               for (int i = 0; i < n[0]; i++) {
                  int g[3] = {fdiv2(b.lo[0] + i), fdiv2(b.lo[1] + j),
                              ndim == 3 ? fdiv2(b.lo[2] + k) : 0};
                  pdst.push_back(cidx(*H, p, i, j, k));
                  psrc.push_back(find_cell(*H, l - 1, g));
               }
         (void)sz;
      }
      if (H->pro_dst[l].upload(pdst) || H->pro_src[l].upload(psrc))
         return hier_fail(H, "upload failed");
   }
   // coarse-fine faces of every fine level, grouped by coarse face so that
   // each coarse face's 2^(ndim-1) fine faces are consecutive in the
   // coarsening order; then the coarse cells next to them
   for (int l = 1; l < nlev; l++) {
      std::vector<CfFace> faces, ffaces;
      struct Key {
         long long cell;
         int d, side;
         long long cface, cm, c0;
         int first;
      };
      std::vector<Key> keys, ikeys;
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         const ampe_box& b = H->boxes[p];
         int n[3] = {box_n(b, 0, ndim), box_n(b, 1, ndim), box_n(b, 2, ndim)};
         long long sy = n[0] + 2, sz = sy * (n[1] + 2);
         for (int d = 0; d < ndim; d++)
            for (int side = 0; side < 2; side++) {
               int t1 = (ndim == 2) ? 1 - d : (d + 1) % 3, t2 = (ndim == 2) ? 2 : (d + 2) % 3;
               // coarse faces along this side
               int cn1 = n[t1] / 2, cn2 = ndim == 3 ? n[t2] / 2 : 1;
               for (int c2 = 0; c2 < cn2; c2++)
                  for (int c1 = 0; c1 < cn1; c1++) {
                     // the coarse cell outside this side
                     int I0[3] = {0, 0, 0};
                     I0[d] = side ? n[d] - 1 : 0;
                     I0[t1] = 2 * c1;
                     if (ndim == 3) I0[t2] = 2 * c2;
                     int gout[3] = {b.lo[0] + I0[0], b.lo[1] + I0[1], ndim == 3 ? b.lo[2] + I0[2] : 0};
                     gout[d] += side ? 1 : -1;
                     bool outside = false;
                     for (int a = 0; a < ndim; a++)
                        if (gout[a] < 0 || gout[a] >= H->dom[l][a]) outside = true;
                     if (outside && !H->periodic) continue;
                     int gf[3] = {gout[0], gout[1], gout[2]};
                     canon(*H, l, gf);
                     // a fine-fine patch boundary: its fine faces still
                     // replace the covered coarse faces in the FAC residual
                     bool fine_fine = find_cell(*H, l, gf) >= 0;
                     std::vector<CfFace>& out = fine_fine ? ffaces : faces;
                     int first = fine_fine ? -1 - (int)ffaces.size() : (int)faces.size();
                     // fine faces of this coarse face, coarsening order
                     for (int r2 = 0; r2 < (ndim == 3 ? 2 : 1); r2++)
                        for (int r1 = 0; r1 < 2; r1++) {
                           int I[3] = {I0[0], I0[1], I0[2]};
                           I[t1] = 2 * c1 + r1;
                           if (ndim == 3) I[t2] = 2 * c2 + r2;
                           int Iq[3] = {I[0], I[1], I[2]};
                           Iq[t1] = I[t1] ^ 1;
                           if (ndim == 3) Iq[t2] = I[t2] ^ 1;
                           int G[3] = {I[0], I[1], I[2]};
                           G[d] += side ? 1 : -1;
                           int F[3] = {I[0], I[1], I[2]};
                           F[d] = side ? n[d] : 0;
                           CfFace f;
                           f.face = fidx(*H, p, d, F[0], F[1], F[2]);
                           long long ci = cidx(*H, p, I[0], I[1], I[2]);
                           long long cg = cidx(*H, p, G[0], G[1], G[2]);
                           f.cm = side ? ci : cg;
                           f.c0 = side ? cg : ci;
                           f.sy = sy;
                           f.sz = sz;
                           int gg[3] = {b.lo[0] + G[0], b.lo[1] + G[1], ndim == 3 ? b.lo[2] + G[2] : 0};
                           canon(*H, l, gg);
                           int par[3] = {fdiv2(gg[0]), fdiv2(gg[1]), ndim == 3 ? fdiv2(gg[2]) : 0};
                           canon(*H, l - 1, par);
                           f.parent = find_cell(*H, l - 1, par);
                           f.mirror = cidx(*H, p, Iq[0], Iq[1], Iq[2]);
                           f.d = d;
                           f.side = side;
                           out.push_back(f);
                        }
                     // the covered coarse cell inside: its face on this side
                     int gin[3] = {fdiv2(b.lo[0] + I0[0]), fdiv2(b.lo[1] + I0[1]),
                                   ndim == 3 ? fdiv2(b.lo[2] + I0[2]) : 0};
                     {
                        int ip = -1;
                        long long icell = find_cell(*H, l - 1, gin, &ip);
                        if (icell < 0) return hier_fail(H, "a level is not nested in the coarser one");
                        const ampe_box& ib = H->boxes[ip];
                        int il[3] = {gin[0] - ib.lo[0], gin[1] - ib.lo[1],
                                     ndim == 3 ? gin[2] - ib.lo[2] : 0};
                        int ifl[3] = {il[0], il[1], il[2]};
                        if (side) ifl[d] += 1;
                        long long ibs = box_n(ib, 0, ndim) + 2;
                        long long ist[3] = {1, ibs, ibs * (box_n(ib, 1, ndim) + 2)};
                        long long inb = side ? icell + ist[d] : icell - ist[d];
                        Key key;
                        key.cell = icell;
                        key.d = d;
                        key.side = side;
                        key.cface = fidx(*H, ip, d, ifl[0], ifl[1], ifl[2]);
                        key.cm = side ? icell : inb;
                        key.c0 = side ? inb : icell;
                        key.first = first;
                        ikeys.push_back(key);
                     }
                     if (fine_fine) continue;
                     // the uncovered coarse cell outside
                     int co[3] = {fdiv2(gf[0]), fdiv2(gf[1]), ndim == 3 ? fdiv2(gf[2]) : 0};
                     int cpatch = -1;
                     long long ccell = find_cell(*H, l - 1, co, &cpatch);
                     if (ccell < 0) return hier_fail(H, "a level is not nested in the coarser one");
                     // the coarse face on ccell's side facing the fine patch
                     const ampe_box& cb = H->boxes[cpatch];
                     int cl[3] = {co[0] - cb.lo[0], co[1] - cb.lo[1], ndim == 3 ? co[2] - cb.lo[2] : 0};
                     int cside = side ? 0 : 1;  // fine patch's hi side -> coarse cell's lo face
                     int cfl[3] = {cl[0], cl[1], cl[2]};
                     if (cside) cfl[d] += 1;
                     long long cface = fidx(*H, cpatch, d, cfl[0], cfl[1], cfl[2]);
                     // neighbour across the face inside the coarse patch's ghost box
                     long long cbs = box_n(cb, 0, ndim) + 2;
                     long long cst[3] = {1, cbs, cbs * (box_n(cb, 1, ndim) + 2)};
                     long long cnb = cside ? ccell + cst[d] : ccell - cst[d];
                     Key key;
                     key.cell = ccell;
                     key.d = d;
                     key.side = cside;
                     key.cface = cface;
                     key.cm = cside ? ccell : cnb;
                     key.c0 = cside ? cnb : ccell;
                     key.first = first;
                     keys.push_back(key);
                  }
            }
      }
      // fine-fine faces follow the coarse-fine ones in the saved fluxes
      for (Key& k : ikeys)
         if (k.first < 0) k.first = (int)faces.size() + (-1 - k.first);
      // corrections of level l-1: one record per coarse cell
      auto records = [&](std::vector<Key>& ks, std::vector<CfCorr>& corr) -> int {
         std::sort(ks.begin(), ks.end(), [](const Key& a, const Key& b) {
            return a.cell != b.cell ? a.cell < b.cell
                                    : (a.d != b.d ? a.d < b.d : a.side < b.side);
         });
         for (size_t i = 0; i < ks.size();) {
            CfCorr c;
            std::memset(&c, 0, sizeof(c));
            c.cell = ks[i].cell;
            int cp = -1;
            for (int q = H->pbeg[l - 1]; q < H->pbeg[l] && cp < 0; q++) {
               const ampe_box& qb = H->boxes[q];
               long long qn = (long long)(box_n(qb, 0, ndim) + 2) * (box_n(qb, 1, ndim) + 2) *
                              (box_n(qb, 2, ndim) + (ndim == 3 ? 2 : 0));
               if (c.cell >= H->coff[q] && c.cell < H->coff[q] + qn) cp = q;
            }
            if (cp < 0) return hier_fail(H, "coarse cell outside its level");
            const ampe_box& cb = H->boxes[cp];
            c.sy = box_n(cb, 0, ndim) + 2;
            c.sz = c.sy * (box_n(cb, 1, ndim) + 2);
            c.nface = 0;
            size_t j = i;
            while (j < ks.size() && ks[j].cell == c.cell) {
               if (c.nface >= 6) return hier_fail(H, "more than six coarse-fine faces on a cell");
               int f = c.nface++;
               c.face[f] = ks[j].cface;
               c.cm[f] = ks[j].cm;
               c.c0[f] = ks[j].c0;
               c.d[f] = ks[j].d;
               c.side[f] = ks[j].side;
               c.fine[f] = ks[j].first;
               j++;
            }
            corr.push_back(c);
            i = j;
         }
         return 0;
      };
      std::vector<CfCorr> corr, icorr;
      if (records(keys, corr) || records(ikeys, icorr)) return -1;
      if (H->cff[l].upload(faces) || H->cfc[l - 1].upload(corr) || H->fff[l].upload(ffaces) ||
          H->cfi[l - 1].upload(icorr))
         return hier_fail(H, "upload failed");
   }

   // weights: cell volume relative to level 0, zero under a finer level
   {
      std::vector<double> w((size_t)H->ncg, 0.0);
      for (int l = 0; l < nlev; l++) {
         double v = 1.0;
         for (int r = 0; r < l; r++) v = v * (ndim == 3 ? 0.125 : 0.25);
         for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
            const ampe_box& b = H->boxes[p];
            int n[3] = {box_n(b, 0, ndim), box_n(b, 1, ndim), box_n(b, 2, ndim)};
            for (int k = 0; k < n[2]; k++)
               for (int j = 0; j < n[1]; j++)
                  for (int i = 0; i < n[0]; i++) {
                     bool cov = false;
                     if (l + 1 < nlev) {
                        int g[3] = {2 * (b.lo[0] + i), 2 * (b.lo[1] + j),
                                    ndim == 3 ? 2 * (b.lo[2] + k) : 0};
                        cov = find_cell(*H, l + 1, g) >= 0;
                     }
                     w[(size_t)cidx(*H, p, i, j, k)] = cov ? 0.0 : v;
                  }
         }
      }
      if (hipMalloc((void**)&H->wt, sizeof(double) * w.size()) != hipSuccess ||
          hipMemcpy(H->wt, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice) !=
              hipSuccess)
         return hier_fail(H, "upload failed");
   }
   H->comp->wt = H->wt;
   *out = H;
   return 0;
}

// ------------------------------------------------------------ kernels
namespace {

template <typename T>
__global__ void __launch_bounds__(256) k_cf_fill(T* __restrict__ f, long long ncg, int ncomp,
                                                 const long long* __restrict__ dst,
                                                 const long long* __restrict__ src,
                                                 const double* __restrict__ w, long long n,
                                                 int mode, int ndim)
{
   long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= n) return;
   const long long* s = src + 9 * e;
   for (int m = 0; m < ncomp; m++) {
      T* a = f + m * ncg;
      T v;
      if (mode == 1) {
         v = a[s[0]];
      } else {
         double x = w[3 * e], y = w[3 * e + 1], z = w[3 * e + 2];
         T q1 = a[s[1]], q2 = a[s[2]], q3 = a[s[3]], q4 = a[s[4]];
         if (ndim == 2) {
            v = (q1 + (q3 - q1) * x) * (1.0 - y) + (q2 + (q4 - q2) * x) * y;
         } else {
            T q5 = a[s[5]], q6 = a[s[6]], q7 = a[s[7]], q8 = a[s[8]];
            v = ((q1 + (q3 - q1) * x) * (1.0 - y) + (q2 + (q4 - q2) * x) * y) * (1.0 - z) +
                ((q5 + (q7 - q5) * x) * (1.0 - y) + (q6 + (q8 - q6) * x) * y) * z;
         }
      }
      a[dst[e]] = v;
   }
}

template <typename T>
__global__ void __launch_bounds__(256) k_coarsen(const T* __restrict__ fsrc,
                                                 T* __restrict__ f, long long ncg, int ncomp,
                                                 const long long* __restrict__ dst,
                                                 const long long* __restrict__ src, long long n,
                                                 double dvf, double dvc, int nchild,
                                                 int quatnorm)
{
   long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= n) return;
   // children in x-fastest order from the first child and the fine strides
   const long long base = src[2 * e], packed = src[2 * e + 1];
   const long long sy = packed & 0xffffffffLL, sz = packed >> 32;
   T acc[4];
   for (int m = 0; m < ncomp; m++) {
      const T* a = fsrc + m * ncg + base;
      T s = 0.0;
      for (int c = 0; c < 8; c++)
         if (c < nchild) s = s + a[(c & 1) + ((c >> 1) & 1) * sy + (c >> 2) * sz] * dvf;
      T v = s / (T)dvc;
      if (m < 4) acc[m] = v;
      if (!quatnorm) f[m * ncg + dst[e]] = v;
   }
   if (quatnorm) {
      double s = 0.0;
      for (int m = 0; m < ncomp; m++) s = s + acc[m] * acc[m];
      double mg = sqrt(s);
      double minv = mg < 1.0e-15 ? 0.0 : 1.0 / mg;
      for (int m = 0; m < ncomp; m++) f[m * ncg + dst[e]] = acc[m] * minv;
   }
}

__global__ void __launch_bounds__(256) k_face_coarsen(const double* __restrict__ src,
                                                      double* __restrict__ dst, long long nside,
                                                      int ncomp, const long long* __restrict__ fd,
                                                      const long long* __restrict__ fs,
                                                      const int* __restrict__ dir, long long n,
                                                      double af0, double af1, double af2,
                                                      double ac0, double ac1, double ac2,
                                                      int nfine)
{
   long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= n) return;
   int d = dir[e];
   double af = d == 0 ? af0 : (d == 1 ? af1 : af2);
   double ac = d == 0 ? ac0 : (d == 1 ? ac1 : ac2);
   for (int m = 0; m < ncomp; m++) {
      const double* a = src + m * nside;
      double acc = 0.0;
      for (int r = 0; r < 4; r++)
         if (r < nfine) acc = acc + a[fs[4 * e + r]] * af;
      dst[m * nside + fd[e]] = acc / ac;
   }
}

template <typename T>
__global__ void __launch_bounds__(256) k_prolong_const_add(T* __restrict__ e, long long ncg,
                                                           int ncomp,
                                                           const long long* __restrict__ dst,
                                                           const long long* __restrict__ src,
                                                           long long n)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   for (int m = 0; m < ncomp; m++) e[m * ncg + dst[t]] = e[m * ncg + dst[t]] + e[m * ncg + src[t]];
}

// interior update u += e over the composite's tiles
template <typename T>
__global__ void __launch_bounds__(256) k_add_interior(LevelArgs a, int ncomp,
                                                      T* __restrict__ u,
                                                      const T* __restrict__ e)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   for (int m = 0; m < ncomp; m++) u[m * a.ncg + c.c0] = u[m * a.ncg + c.c0] + e[m * a.ncg + c.c0];
}

// per-tile sum of w r^2 over ncomp components (weights of interior cells)
template <typename T>
__global__ void __launch_bounds__(256) k_wsumsq(LevelArgs a, int ncomp,
                                                const T* __restrict__ r,
                                                const double* __restrict__ w,
                                                double* __restrict__ partial)
{
   CellPos c;
   double v = 0.0, sw = 0.0;
   if (cell_pos(a, 0, c)) {
      double wc = w[c.c0];
      for (int m = 0; m < ncomp; m++) {
         double x = (double)r[m * a.ncg + c.c0];
         v += wc * x * x;
         sw += wc;
      }
   }
   __shared__ double sh[2][TX * TY];
   int t = threadIdx.x + TX * threadIdx.y;
   sh[0][t] = v;
   sh[1][t] = sw;
   __syncthreads();
   for (int s = TX * TY / 2; s > 0; s >>= 1) {
      if (t < s) {
         sh[0][t] = sh[0][t] + sh[0][t + s];
         sh[1][t] = sh[1][t] + sh[1][t + s];
      }
      __syncthreads();
   }
   if (t == 0) {
      partial[2 * blockIdx.x] = sh[0][0];
      partial[2 * blockIdx.x + 1] = sh[1][0];
   }
}

// Ewing's fix (efo_ewingfixflux*, source/fortran/2d/ellipticfacops.m4:
// 543-780) as an effective ghost: the flux D (u_I - u_g*) / h across the
// face equals longwt D (u_I - u_g) / h + tranwt D (u_I' - u_I) / h on a lower
// side (the mirror image on an upper side), with u_g the constant-refined
// coarse value kept in gsave; longwt = 2 tranwt, tranwt = 1.0/(1+ratio) in
// single precision as the reference writes it.
template <typename T>
__global__ void __launch_bounds__(256) k_ewing_ghosts(T* __restrict__ u,
                                                      const CfFace* __restrict__ ff, long long n,
                                                      const T* __restrict__ gsave,
                                                      T tranwt, T longwt)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfFace f = ff[t];
   long long I = f.side ? f.cm : f.c0, g = f.side ? f.c0 : f.cm;
   T uI = u[I];
   u[g] = uI - longwt * (uI - gsave[t]) - tranwt * (u[f.mirror] - uI);
}

// gsave[t] = coarse parent value of face t's ghost (CONSTANT_REFINE); with
// write_ghost the ghost itself takes it too
template <typename T>
__global__ void __launch_bounds__(256) k_cf_save(T* __restrict__ u,
                                                 const CfFace* __restrict__ ff, long long n,
                                                 T* __restrict__ gsave, int write_ghost)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfFace f = ff[t];
   T v = u[f.parent];
   gsave[t] = v;
   if (write_ghost) u[f.side ? f.c0 : f.cm] = v;
}

// fine coarse-fine face fluxes D (w_c0 - w_cm) / h (w = u, or S u for the
// orientation block), saved for the coarser level's composite residual
template <typename T>
__global__ void __launch_bounds__(256) k_oflux_save(const T* __restrict__ u,
                                                    const T* __restrict__ S,
                                                    const T* __restrict__ D, T Dc,
                                                    const CfFace* __restrict__ ff, long long n,
                                                    T di0, T di1, T di2,
                                                    T* __restrict__ oflux)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfFace f = ff[t];
   T dinv = f.d == 0 ? di0 : (f.d == 1 ? di1 : di2);
   T dd = D ? D[f.face] : Dc;
   T wc = S ? S[f.c0] * u[f.c0] : u[f.c0];
   T wm = S ? S[f.cm] * u[f.cm] : u[f.cm];
   oflux[t] = dinv * dd * (wc - wm);
}

// composite residual: the coarse flux on each coarse-fine face replaced by
// the coarsened saved fine fluxes
template <typename T>
__global__ void __launch_bounds__(256) k_fac_corr(T* __restrict__ res,
                                                  const T* __restrict__ u,
                                                  const T* __restrict__ S,
                                                  const T* __restrict__ D, T Dc,
                                                  const T* __restrict__ M, T Mc,
                                                  T gamma, int kind,
                                                  const CfCorr* __restrict__ cc, long long n,
                                                  const T* __restrict__ oflux, int nfine,
                                                  T di0, T di1, T di2, T af, T ac)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const CfCorr& e = cc[t];
   T delta = 0.0;
   for (int f = 0; f < e.nface; f++) {
      int d = e.d[f];
      T dinv = d == 0 ? di0 : (d == 1 ? di1 : di2);
      T dd = D ? D[e.face[f]] : Dc;
      T wc = S ? S[e.c0[f]] * u[e.c0[f]] : u[e.c0[f]];
      T wm = S ? S[e.cm[f]] * u[e.cm[f]] : u[e.cm[f]];
      T fold = dinv * dd * (wc - wm);
      T acc = 0.0;
      for (int r = 0; r < nfine; r++) acc = acc + oflux[e.fine[f] + r] * af;
      T fnew = acc / ac;
      T t1 = dinv * (fnew - fold);
      delta = delta + (e.side[f] ? t1 : -t1);
   }
   T m = kind == 1 ? gamma * S[e.cell] : (M ? M[e.cell] : Mc);
   res[e.cell] = res[e.cell] - m * delta;
}

}  // namespace

static dim3 grid1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// ------------------------------------------------------------ transfers
int hier_ghost_fill_level(const Hier& H, int l, double* f, int ncomp, int cf, hipStream_t s)
{
   if (ghost_fill(*H.lev[l], f, ncomp, s)) return -1;
   if (l > 0 && cf > 0 && H.cf_dst[l].n > 0) {
      hipLaunchKernelGGL(k_cf_fill<double>, grid1(H.cf_dst[l].n), dim3(256), 0, s, f, H.ncg, ncomp,
                         H.cf_dst[l].p, H.cf_src[l].p, H.cf_w[l].p, H.cf_dst[l].n, cf, H.ndim);
      AMPE_HIP_CHECK(hipGetLastError());
   }
   return 0;
}

int hier_ghost_fill(const Hier& H, double* f, int ncomp, int cf, hipStream_t s)
{
   for (int l = 0; l < H.nlev; l++)
      if (hier_ghost_fill_level(H, l, f, ncomp, cf, s)) return -1;
   return 0;
}

static double cell_vol(const Hier& H, int l)
{
   return H.dx[l][0] * H.dx[l][1] * (H.ndim == 3 ? H.dx[l][2] : 1.0);
}

int hier_coarsen_level(const Hier& H, int l, double* f, int ncomp, int quatnorm, hipStream_t s)
{
   if (l + 1 >= H.nlev || H.crs_dst[l].n == 0) return 0;
   if (quatnorm && ncomp > 4) {
      set_error("hier_coarsen: at most four components with the normalisation");
      return -1;
   }
   hipLaunchKernelGGL(k_coarsen<double>, grid1(H.crs_dst[l].n), dim3(256), 0, s, f, f, H.ncg, ncomp,
                      H.crs_dst[l].p, H.crs_src[l].p, H.crs_dst[l].n, cell_vol(H, l + 1),
                      cell_vol(H, l), H.ndim == 3 ? 8 : 4, quatnorm);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int hier_coarsen(const Hier& H, double* f, int ncomp, int quatnorm, hipStream_t s)
{
   for (int l = H.nlev - 2; l >= 0; l--)
      if (hier_coarsen_level(H, l, f, ncomp, quatnorm, s)) return -1;
   return 0;
}

static void face_areas(const Hier& H, int l, double af[3], double ac[3])
{
   for (int d = 0; d < 3; d++) {
      af[d] = 1.0;
      ac[d] = 1.0;
      for (int a = 0; a < H.ndim; a++)
         if (a != d) {
            af[d] = af[d] * H.dx[l + 1][a];
            ac[d] = ac[d] * H.dx[l][a];
         }
   }
}

int hier_coarsen_faces_level(const Hier& H, int l, const double* src, double* dst, int ncomp,
                             hipStream_t s)
{
   if (l + 1 >= H.nlev || H.fcrs_dst[l].n == 0) return 0;
   double af[3], ac[3];
   face_areas(H, l, af, ac);
   hipLaunchKernelGGL(k_face_coarsen, grid1(H.fcrs_dst[l].n), dim3(256), 0, s, src, dst,
                      H.nside_tot, ncomp, H.fcrs_dst[l].p, H.fcrs_src[l].p, H.fcrs_dir[l].p,
                      H.fcrs_dst[l].n, af[0], af[1], af[2], ac[0], ac[1], ac[2],
                      H.ndim == 3 ? 4 : 2);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

static ampe_model_params lev_params(const Hier& H, const ampe_model_params& p0, int l)
{
   ampe_model_params p = p0;
   for (int a = 0; a < 3; a++) p.dx[a] = H.dx[l][a];
   return p;
}

// the face-area ratio of the fine faces on a coarse face (same in every
// direction for isotropic refinement): per direction d the areas of d
static void corr_areas(const Hier& H, int l, double& af, double& ac)
{
   double a[3], c[3];
   face_areas(H, l, a, c);
   af = a[0];
   ac = c[0];
}

// ------------------------------------------------------------ composite RHS
int hier_phase_rhs(const Hier& H, const ampe_model_params& p0, const double* phi,
                   const double* q, const double* cl, const double* cs, double* rhs,
                   hipStream_t s)
{
   for (int l = H.nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      if (launch_phase_rhs(*H.lev[l], p, phi, q, cl, cs, rhs, s)) return -1;
      if (l + 1 < H.nlev) {
         double af, ac;
         corr_areas(H, l, af, ac);
         ampe_model_params pf = lev_params(H, p0, l + 1);
         if (launch_hier_phase_corr(H.ndim, H.cfc[l].p, (int)H.cfc[l].n, H.cff[l + 1].p, p, pf,
                                    H.ncg, phi, q, cl, cs, af, ac, rhs, s))
            return -1;
      }
   }
   return 0;
}

int hier_quat_face_norm2(const Hier& H, const ampe_model_params& p0, const double* q,
                         double* n2, hipStream_t s)
{
   for (int l = 0; l < H.nlev; l++)
      if (launch_quat_grad(*H.lev[l], lev_params(H, p0, l), q, nullptr, n2, s)) return -1;
   return 0;
}

int hier_quat_rhs(const Hier& H, const ampe_model_params& p0, const double* phi, const double* q,
                  const double* n2, double* fc, double* rhs, hipStream_t s)
{
   for (int l = H.nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      if (launch_quat_face_coef(*H.lev[l], p, phi, n2, fc, s)) return -1;
      if (launch_quat_rhs_fc(*H.lev[l], p, phi, q, fc, rhs, s)) return -1;
      if (l + 1 < H.nlev) {
         double af, ac;
         corr_areas(H, l, af, ac);
         if (launch_hier_quat_corr(H.ndim, H.cfc[l].p, (int)H.cfc[l].n, H.cff[l + 1].p, p,
                                   lev_params(H, p0, l + 1), H.ncg, phi, q, fc, af, ac, rhs, s))
            return -1;
      }
   }
   return 0;
}

int hier_conc_coeffs(const Hier& H, const ampe_model_params& p0, const double* phi,
                     const double* cl, const double* cs, double* Dc, double* Dphi, hipStream_t s)
{
   for (int l = 0; l < H.nlev; l++)
      if (launch_conc_coeffs(*H.lev[l], lev_params(H, p0, l), phi, cl, cs, Dc, Dphi, s))
         return -1;
   for (int l = H.nlev - 2; l >= 0; l--) {
      if (hier_coarsen_faces_level(H, l, Dc, Dc, 1, s)) return -1;
      if (hier_coarsen_faces_level(H, l, Dphi, Dphi, 1, s)) return -1;
   }
   return 0;
}

int hier_conc_rhs(const Hier& H, const ampe_model_params& p0, const double* c, const double* phi,
                  const double* Dc, const double* Dphi, double* rhs, hipStream_t s)
{
   for (int l = H.nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      if (launch_conc_rhs(*H.lev[l], p, c, phi, Dc, Dphi, rhs, s)) return -1;
      if (l + 1 < H.nlev) {
         double af, ac;
         corr_areas(H, l, af, ac);
         if (launch_hier_conc_corr(H.ndim, H.cfc[l].p, (int)H.cfc[l].n, H.cff[l + 1].p, p,
                                   lev_params(H, p0, l + 1), c, phi, Dc, Dphi, af, ac, rhs, s))
            return -1;
      }
   }
   return 0;
}

// The three-phase multi-order model on the hierarchy (phase concentrations
// ph6 of every cell already set): finest to coarsest, each level's fused
// kernels, then the coarse-fine corrections with the finer level's fluxes;
// the face diffusivities of the finer levels are coarsened onto the coarser
// faces first, as for the KKS coefficients.
int hier_phase_rhs_tern3(const Hier& H, const ampe_model_params& p0, int norder,
                         const double* phi, const double* ph6, double* rhs, hipStream_t s)
{
   for (int l = H.nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      if (launch_phase_rhs_tern3(*H.lev[l], p, norder, phi, ph6, nullptr, rhs, s)) return -1;
      if (l + 1 < H.nlev) {
         double af, ac;
         corr_areas(H, l, af, ac);
         if (launch_hier_tern3_phase_corr(H.ndim, H.cfc[l].p, (int)H.cfc[l].n, H.cff[l + 1].p,
                                          p, lev_params(H, p0, l + 1), norder, H.ncg, phi, af,
                                          ac, rhs, s))
            return -1;
      }
   }
   return 0;
}

// the multi-order model without composition (grain growth): the same
// per-order gradient flux as the three-phase model, so the same correction
int hier_phase_rhs_multiorder(const Hier& H, const ampe_model_params& p0, int norder,
                              const double* phi, double* rhs, hipStream_t s)
{
   for (int l = H.nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      if (launch_phase_rhs_multiorder(*H.lev[l], p, norder, phi, nullptr, rhs, s)) return -1;
      if (l + 1 < H.nlev) {
         double af, ac;
         corr_areas(H, l, af, ac);
         if (launch_hier_tern3_phase_corr(H.ndim, H.cfc[l].p, (int)H.cfc[l].n, H.cff[l + 1].p,
                                          p, lev_params(H, p0, l + 1), norder, H.ncg, phi, af,
                                          ac, rhs, s))
            return -1;
      }
   }
   return 0;
}

int hier_ebs3_coeffs(const Hier& H, const ampe_model_params& p0, int norder, const double* phi,
                     double* DL, double* DA, double* DB, hipStream_t s)
{
   for (int l = 0; l < H.nlev; l++)
      if (launch_ebs3_coeffs(*H.lev[l], lev_params(H, p0, l), norder, phi, DL, DA, DB, s))
         return -1;
   for (int l = H.nlev - 2; l >= 0; l--) {
      if (hier_coarsen_faces_level(H, l, DL, DL, 1, s)) return -1;
      if (hier_coarsen_faces_level(H, l, DA, DA, 1, s)) return -1;
      if (hier_coarsen_faces_level(H, l, DB, DB, 1, s)) return -1;
   }
   return 0;
}

int hier_conc_rhs_ebs3(const Hier& H, const ampe_model_params& p0, const double* ph6,
                       const double* DL, const double* DA, const double* DB, double* rhs,
                       hipStream_t s)
{
   for (int l = H.nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      if (launch_conc_rhs_ebs3(*H.lev[l], p, ph6, DL, DA, DB, rhs, s)) return -1;
      if (l + 1 < H.nlev) {
         double af, ac;
         corr_areas(H, l, af, ac);
         if (launch_hier_ebs3_corr(H.ndim, H.cfc[l].p, (int)H.cfc[l].n, H.cff[l + 1].p, p,
                                   lev_params(H, p0, l + 1), H.ncg, ph6, DL, DA, DB, af, ac,
                                   rhs, s))
            return -1;
      }
   }
   return 0;
}

// ------------------------------------------------------------ FAC
// The FAC work arrays of one element type (FP64 for the block solves of the
// FP64 preconditioner, FP32 for the FP32 one), grown on demand.
template <typename T>
static FacWork<T>& fac_work(Hier& H);
template <>
FacWork<double>& fac_work<double>(Hier& H)
{
   return H.wd;
}
template <>
FacWork<float>& fac_work<float>(Hier& H)
{
   return H.wf;
}

template <typename T>
static int fac_grow(Hier& H, int ncomp)
{
   FacWork<T>& W = fac_work<T>(H);
   long long need = H.ncg * ncomp;
   if (need <= W.cap) return 0;
   W.release();
   long long nface = 0;
   for (int l = 0; l < H.nlev; l++) nface = std::max(nface, H.cff[l].n + H.fff[l].n);
   AMPE_HIP_CHECK(hipMalloc((void**)&W.e, sizeof(T) * need));
   AMPE_HIP_CHECK(hipMalloc((void**)&W.r, sizeof(T) * need));
   AMPE_HIP_CHECK(hipMalloc((void**)&W.tmp, sizeof(T) * need));
   // saved fine fluxes per level and component
   AMPE_HIP_CHECK(
       hipMalloc((void**)&W.oflux, sizeof(T) * (size_t)(nface + 1) * kMaxLev * ncomp));
   AMPE_HIP_CHECK(hipMalloc((void**)&W.gsave, sizeof(T) * (size_t)(nface + 1)));
   W.cap = need;
   return 0;
}

template <typename T>
struct FacCtx {
   Hier* H;
   const HBlock<T>* B;
   const ampe_fac_params* P;
   int ncomp;
   long long oflux_stride;  // per (level, component) slot
   hipStream_t s;
   FacWork<T>* W;
};

template <typename T>
static T* oflux_of(const FacCtx<T>& X, int l, int m)
{
   return X.W->oflux + ((long long)l * X.ncomp + m) * X.oflux_stride;
}

template <typename T>
static const T* comp_ptr(const FacCtx<T>& X, const T* f, int m)
{
   return f + (long long)m * X.H->ncg;
}

static double tranwt_ref() { return (double)(1.0f / (float)(1 + 2)); }

template <typename T>
static EllOp<T> ell_op(const HBlock<T>& B)
{
   EllOp<T> op;
   op.D = B.D;
   op.Dc = B.D_const;
   op.C = B.C;
   op.Cc = B.C_const;
   op.M = B.M;
   op.Mc = B.M_const;
   return op;
}

// coarse-fine ghosts of level l (> 0) for component array u: constant refine
// saved per face; the elliptic blocks then take Ewing's effective ghosts
template <typename T>
static int fac_cf_ghosts(const FacCtx<T>& X, int l, T* u, bool refresh)
{
   const Hier& H = *X.H;
   if (l == 0) return 0;
   long long n = H.cff[l].n;
   bool ewing = X.B->kind == 0 && X.B->ewing;
   if (refresh) {
      // the full coarse-fine fill by CONSTANT_REFINE of every ghost (corner
      // ghosts included), then the per-face saved values
      if (H.cf_dst[l].n > 0) {
         hipLaunchKernelGGL(k_cf_fill<T>, grid1(H.cf_dst[l].n), dim3(256), 0, X.s, u, H.ncg, 1,
                            H.cf_dst[l].p, H.cf_src[l].p, H.cf_w[l].p, H.cf_dst[l].n, 1, H.ndim);
         AMPE_HIP_CHECK(hipGetLastError());
      }
      if (n > 0 && ewing) {
         hipLaunchKernelGGL(k_cf_save<T>, grid1(n), dim3(256), 0, X.s, u, H.cff[l].p, n,
                            X.W->gsave, 0);
         AMPE_HIP_CHECK(hipGetLastError());
      }
   }
   if (n > 0 && ewing) {
      double tw = tranwt_ref();
      hipLaunchKernelGGL(k_ewing_ghosts<T>, grid1(n), dim3(256), 0, X.s, u, H.cff[l].p, n,
                         X.W->gsave, (T)tw, (T)(2.0 * tw));
      AMPE_HIP_CHECK(hipGetLastError());
   }
   return 0;
}

// The Ewing effective ghosts depend on the component being smoothed; with
// several components (orientation block, no Ewing) the saved values are not
// used.
template <typename T>
static int fac_colour(const FacCtx<T>& X, int l, T* u, const T* f, int color)
{
   const Hier& H = *X.H;
   const Level& L = *H.lev[l];
   const HBlock<T>& B = *X.B;
   if (ghost_fill(L, u, 1, X.s)) return -1;
   if (l > 0 && fac_cf_ghosts(X, l, u, false)) return -1;
   if (B.kind == 1)
      return launch_quat_rbgs<T>(L, B.D, B.S, B.gamma, H.dx[l], 1, f, u, color, X.s);
   return launch_rbgs<T>(L, ell_op(B), H.dx[l], f, u, color, false, X.s);
}

template <typename T>
static int fac_smooth(const FacCtx<T>& X, int l, const T* f, T* u, int sweeps)
{
   const Hier& H = *X.H;
   for (int m = 0; m < X.ncomp; m++) {
      T* um = u + (long long)m * H.ncg;
      const T* fm = comp_ptr(X, f, m);
      if (ghost_fill(*H.lev[l], um, 1, X.s)) return -1;
      if (l > 0 && fac_cf_ghosts(X, l, um, true)) return -1;
      for (int sw = 0; sw < sweeps; sw++) {
         if (fac_colour(X, l, um, fm, 0)) return -1;
         if (fac_colour(X, l, um, fm, 1)) return -1;
         if (ghost_fill(*H.lev[l], um, 1, X.s)) return -1;
      }
   }
   return 0;
}

// composite residual on level l (EllipticFACOps::computeCompositeResidualOnLevel)
template <typename T>
static int fac_comp_residual(const FacCtx<T>& X, int l, T* res, T* u, const T* f)
{
   const Hier& H = *X.H;
   const HBlock<T>& B = *X.B;
   const Level& L = *H.lev[l];
   for (int m = 0; m < X.ncomp; m++) {
      T* um = u + (long long)m * H.ncg;
      const T* fm = comp_ptr(X, f, m);
      T* rm = res + (long long)m * H.ncg;
      if (ghost_fill(L, um, 1, X.s)) return -1;
      if (l > 0 && fac_cf_ghosts(X, l, um, true)) return -1;
      if (B.kind == 1) {
         if (launch_quat_residual<T>(L, B.D, B.S, B.gamma, H.dx[l], 1, fm, um, rm, nullptr, X.s))
            return -1;
      } else if (launch_residual<T>(L, ell_op(B), H.dx[l], fm, um, rm, nullptr, X.s)) {
         return -1;
      }
      const T di[3] = {(T)(1.0 / H.dx[l][0]), (T)(1.0 / H.dx[l][1]), (T)(1.0 / H.dx[l][2])};
      // every coarse face on a fine patch boundary takes the coarsened saved
      // fine fluxes: the uncovered cells next to coarse-fine faces, then the
      // covered cells inside the fine patches
      for (int t = 0; t < 2 && l + 1 < H.nlev; t++) {
         const DevVec<CfCorr>& tab = t == 0 ? H.cfc[l] : H.cfi[l];
         if (tab.n == 0) continue;
         double af, ac;
         corr_areas(H, l, af, ac);
         hipLaunchKernelGGL(k_fac_corr<T>, grid1(tab.n), dim3(256), 0, X.s, rm, um,
                            B.kind == 1 ? B.S : nullptr, B.D, B.D_const, B.M, B.M_const, B.gamma,
                            B.kind, tab.p, tab.n, oflux_of(X, l + 1, m), H.ndim == 3 ? 4 : 2,
                            di[0], di[1], di[2], (T)af, (T)ac);
         AMPE_HIP_CHECK(hipGetLastError());
      }
      for (int t = 0; t < 2 && l > 0; t++) {
         const DevVec<CfFace>& tab = t == 0 ? H.cff[l] : H.fff[l];
         if (tab.n == 0) continue;
         hipLaunchKernelGGL(k_oflux_save<T>, grid1(tab.n), dim3(256), 0, X.s, um,
                            B.kind == 1 ? B.S : nullptr, B.D, B.D_const, tab.p, tab.n, di[0],
                            di[1], di[2], oflux_of(X, l, m) + (t == 0 ? 0 : H.cff[l].n));
         AMPE_HIP_CHECK(hipGetLastError());
      }
   }
   return 0;
}

template <typename T>
static int fac_norm(const FacCtx<T>& X, const T* r, double* out)
{
   const Hier& H = *X.H;
   const Level& C = *H.comp;
   LevelArgs a = level_args(C);
   hipLaunchKernelGGL(k_wsumsq<T>, dim3(C.ncell_tiles), dim3(TX, TY), 0, X.s, a, X.ncomp, r,
                      H.wt, C.partial);
   AMPE_HIP_CHECK(hipGetLastError());
   double h[2];
   if (reduce_partials(C, C.ncell_tiles, 2, 0, h, X.s)) return -1;
   *out = std::sqrt(h[0] / h[1]);
   return 0;
}

template <typename T>
static int coarsen_level_t(const Hier& H, int l, const T* src, T* dst, int ncomp, hipStream_t s)
{
   if (l + 1 >= H.nlev || H.crs_dst[l].n == 0) return 0;
   hipLaunchKernelGGL(k_coarsen<T>, grid1(H.crs_dst[l].n), dim3(256), 0, s, src, dst, H.ncg,
                      ncomp, H.crs_dst[l].p, H.crs_src[l].p, H.crs_dst[l].n, cell_vol(H, l + 1),
                      cell_vol(H, l), H.ndim == 3 ? 8 : 4, 0);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
static int full_comp_residual(const FacCtx<T>& X, T* r, T* u, const T* f, double* norm)
{
   const Hier& H = *X.H;
   int lmax = H.nlev - 1;
   if (fac_comp_residual(X, lmax, r, u, f)) return -1;
   for (int l = lmax - 1; l >= 0; l--) {
      if (coarsen_level_t<T>(H, l, u, u, X.ncomp, X.s)) return -1;
      if (fac_comp_residual(X, l, r, u, f)) return -1;
      if (coarsen_level_t<T>(H, l, r, r, X.ncomp, X.s)) return -1;
   }
   return norm ? fac_norm(X, r, norm) : 0;
}

template <typename T>
static int fac_cycle(const FacCtx<T>& X, T* e, T* r, int l)
{
   const Hier& H = *X.H;
   const ampe_fac_params& P = *X.P;
   if (l == 0) {
      // solveCoarsestLevel, "redblack": sweeps with the tolerance test
      const HBlock<T>& B = *X.B;
      for (int m = 0; m < X.ncomp; m++) {
         T* em = e + (long long)m * H.ncg;
         const T* rm = comp_ptr(X, r, m);
         if (ghost_fill(*H.lev[0], em, 1, X.s)) return -1;
         for (int sw = 0; sw < P.coarse_its; sw++) {
            bool want = P.coarse_tol >= 0.0;
            double rmax = 0.0, bmax = 0.0;
            for (int color = 0; color < 2; color++) {
               if (ghost_fill(*H.lev[0], em, 1, X.s)) return -1;
               if (B.kind == 1) {
                  if (launch_quat_rbgs<T>(*H.lev[0], B.D, B.S, B.gamma, H.dx[0], 1, rm, em,
                                          color, X.s))
                     return -1;
               } else {
                  if (launch_rbgs<T>(*H.lev[0], ell_op(B), H.dx[0], rm, em, color, want, X.s))
                     return -1;
                  if (want) {
                     double mx;
                     if (reduce_partials(*H.lev[0], H.lev[0]->ncell_tiles, 1, 1, &mx, X.s))
                        return -1;
                     (color ? bmax : rmax) = mx;
                  }
               }
            }
            if (ghost_fill(*H.lev[0], em, 1, X.s)) return -1;
            if (want && B.kind == 0 && std::max(rmax, bmax) <= P.coarse_tol) break;
         }
         if (ghost_fill(*H.lev[0], em, 1, X.s)) return -1;
      }
      return 0;
   }
   if (fac_smooth(X, l, r, e, P.pre)) return -1;
   if (fac_comp_residual(X, l, X.W->tmp, e, r)) return -1;
   // restrictResidual: the covered cells of r on level l-1 take the
   // conservative average of the level-l composite residual in tmp
   if (coarsen_level_t<T>(H, l - 1, X.W->tmp, r, X.ncomp, X.s)) return -1;
   if (fac_cycle(X, e, r, l - 1)) return -1;
   if (H.pro_dst[l].n > 0) {
      hipLaunchKernelGGL(k_prolong_const_add<T>, grid1(H.pro_dst[l].n), dim3(256), 0, X.s, e,
                         H.ncg, X.ncomp, H.pro_dst[l].p, H.pro_src[l].p, H.pro_dst[l].n);
      AMPE_HIP_CHECK(hipGetLastError());
   }
   return fac_smooth(X, l, r, e, P.post);
}

// FACPreconditioner::solveSystem on the block B: cycles of corrections to u
// from the composite residual against f, all in the element type T (FP64
// for the FP64 preconditioner, FP32 for the FP32 one).
template <typename T>
static int fac_solve_t(Hier& H, const HBlock<T>& B, const ampe_fac_params& P, const T* f, T* u,
                       int ncomp, double tol, double* norm, int* cycles, hipStream_t s)
{
   if (ncomp < 1 || (B.kind == 0 && ncomp != 1) || (B.kind == 1 && !B.S)) {
      set_error("hier_fac_solve: one component for an elliptic block; S for the orientation");
      return -1;
   }
   if (fac_grow<T>(H, ncomp)) return -1;
   FacWork<T>& W = fac_work<T>(H);
   long long nface = 0;
   for (int l = 0; l < H.nlev; l++) nface = std::max(nface, H.cff[l].n + H.fff[l].n);
   FacCtx<T> X{&H, &B, &P, ncomp, nface + 1, s, &W};
   long long n = H.ncg * ncomp;
   if (fill_zero<T>(W.r, n, s)) return -1;
   // the initial residual drives the first cycle; its norm is not needed
   // (the cycle loop tests the norm after each cycle)
   double nrm = 0.0;
   if (full_comp_residual(X, W.r, u, f, nullptr)) return -1;
   int it = 0;
   LevelArgs a = level_args(*H.comp);
   do {
      if (fill_zero<T>(W.e, n, s)) return -1;
      if (fac_cycle(X, W.e, W.r, H.nlev - 1)) return -1;
      hipLaunchKernelGGL(k_add_interior<T>, dim3(H.comp->ncell_tiles), dim3(TX, TY), 0, s, a,
                         ncomp, u, W.e);
      AMPE_HIP_CHECK(hipGetLastError());
      for (int l = H.nlev - 2; l >= 0; l--)
         if (coarsen_level_t<T>(H, l, u, u, ncomp, s)) return -1;
      if (full_comp_residual(X, W.r, u, f, &nrm)) return -1;
      it++;
   } while (nrm > tol && it < P.max_cycles);
   if (norm) *norm = nrm;
   if (cycles) *cycles = it;
   return nrm < tol ? 0 : 1;
}

static HBlock<double> hblock_of(const ampe_hier_block& B)
{
   return HBlock<double>{B.kind,  B.D, B.D_const, B.C,     B.C_const,
                         B.M,     B.M_const, B.S, B.gamma, B.ewing};
}

int hier_fac_solve(Hier& H, const ampe_hier_block& B, const ampe_fac_params& P, const double* f,
                   double* u, int ncomp, double tol, double* norm, int* cycles, hipStream_t s)
{
   HBlock<double> b = hblock_of(B);
   return fac_solve_t<double>(H, b, P, f, u, ncomp, tol, norm, cycles, s);
}

int hier_fac_solve_f32(Hier& H, const HBlock<float>& B, const ampe_fac_params& P, const float* f,
                       float* u, int ncomp, double tol, double* norm, int* cycles, hipStream_t s)
{
   return fac_solve_t<float>(H, B, P, f, u, ncomp, tol, norm, cycles, s);
}

Level* hier_composite(Hier* H) { return H->comp; }

int hier_nlev(const Hier* H) { return H->nlev; }

int hier_weight_sum(Hier* H, double* sum)
{
   std::vector<double> w((size_t)H->ncg);
   AMPE_HIP_CHECK(hipMemcpy(w.data(), H->wt, sizeof(double) * w.size(), hipMemcpyDeviceToHost));
   double s = 0.0;
   for (double v : w) s += v;
   *sum = s;
   return 0;
}

int hier_sim_sync_fill(Hier* H, double* y, int nc, int nord, int ic, int ncs, int iq, int qlen,
                       hipStream_t s)
{
   if (hier_coarsen(*H, y, nord, 0, s)) return -1;
   if (ic >= 0 && hier_coarsen(*H, y + (long long)ic * H->ncg, ncs, 0, s)) return -1;
   if (iq >= 0 && hier_coarsen(*H, y + (long long)iq * H->ncg, qlen, 1, s)) return -1;
   return hier_ghost_fill(*H, y, nc, 2, s);
}

int hier_precond_coeffs(const Hier& H, const ampe_model_params& p0, double gamma,
                        const double* phi, const double* cl, const double* cs, const double* n2,
                        double* Mphi, double* Cphi, double* fc, double* sqrtm, hipStream_t s)
{
   bool kim = p0.phi_mobility_type == AMPE_PHI_MOB_KIM;
   for (int l = 0; l < H.nlev; l++) {
      ampe_model_params p = lev_params(H, p0, l);
      const Level& L = *H.lev[l];
      if (kim && launch_phase_mobility(L, p, cl, cs, Mphi, s)) return -1;
      if (launch_phase_C(L, p, gamma, phi, kim ? Mphi : nullptr, Cphi, s)) return -1;
      if (fc && launch_quat_face_coef(L, p, phi, n2, fc, s)) return -1;
   }
   // sqrt(M_q) of every ghosted entry (cell-wise, no spacing)
   if (sqrtm && launch_sqrt_quat_mobility(*H.comp, p0, phi, sqrtm, s)) return -1;
   return 0;
}

// ------------------------------------------------------------ regrid
namespace {

__global__ void __launch_bounds__(256) k_copy_cells(const double* __restrict__ fo, long long ncg_o,
                                                    double* __restrict__ fn, long long ncg_n,
                                                    int ncomp, const long long* __restrict__ dst,
                                                    const long long* __restrict__ src, long long n)
{
   long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= n) return;
   for (int m = 0; m < ncomp; m++) fn[m * ncg_n + dst[e]] = fo[m * ncg_o + src[e]];
}

bool same_domain(const Hier& A, const Hier& B)
{
   if (A.ndim != B.ndim || A.periodic != B.periodic) return false;
   for (int a = 0; a < 3; a++)
      if (A.dom[0][a] != B.dom[0][a] || A.dx[0][a] != B.dx[0][a]) return false;
   return true;
}

}  // namespace

// The regrid fill of QuatModel::Regrid (source/QuatModel.cc:2733-2790; the
// refine schedule of initializeLevelData, :2997-3004): a cell of the new
// hierarchy's level l takes the old level l's value where an old patch holds
// it, else LINEAR_REFINE of the new level l-1 (the coarse-fine ghost
// interpolation), coarsest level first.  Interior cells only.
int hier_transfer(const Hier& Ho, const Hier& Hn, const double* fo, double* fn, int ncomp,
                  hipStream_t s)
{
   if (!same_domain(Ho, Hn) || ncomp < 1) {
      set_error("hier_transfer: the hierarchies must share dimension, level-0 domain, spacing "
                "and periodicity");
      return -1;
   }
   const int ndim = Hn.ndim;
   for (int l = 0; l < Hn.nlev; l++) {
      std::vector<long long> cd, cs, rd, rs;
      std::vector<double> rw;
      for (int p = Hn.pbeg[l]; p < Hn.pbeg[l + 1]; p++) {
         const ampe_box& b = Hn.boxes[p];
         int n[3] = {box_n(b, 0, ndim), box_n(b, 1, ndim), box_n(b, 2, ndim)};
         for (int k = 0; k < n[2]; k++)
            for (int j = 0; j < n[1]; j++)
               for (int i = 0; i < n[0]; i++) {
                  int g[3] = {b.lo[0] + i, b.lo[1] + j, ndim == 3 ? b.lo[2] + k : 0};
                  long long dst = cidx(Hn, p, i, j, k);
                  long long src = l < Ho.nlev ? find_cell(Ho, l, g) : -1;
                  if (src >= 0) {
                     cd.push_back(dst);
                     cs.push_back(src);
                     continue;
                  }
                  if (l == 0) {
                     set_error("hier_transfer: the old level 0 does not cover the new one");
                     return -1;
                  }
                  int cp[3] = {fdiv2(g[0]), fdiv2(g[1]), ndim == 3 ? fdiv2(g[2]) : 0};
                  canon(Hn, l - 1, cp);
                  long long sc = find_cell(Hn, l - 1, cp);
                  if (sc < 0) {
                     set_error("hier_transfer: a level is not nested in the coarser one");
                     return -1;
                  }
                  rd.push_back(dst);
                  rs.push_back(sc);
                  int gc[3] = {0, 0, 0};
                  double w[3] = {0.0, 0.0, 0.0};
                  for (int a = 0; a < ndim; a++)
                     lin_index(g[a], Hn.dx[l][a], Hn.dx[l - 1][a], gc[a], w[a]);
                  for (int a = 0; a < 3; a++) rw.push_back(w[a]);
                  static const int off[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0},
                                                {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
                  for (int c = 0; c < 8; c++) {
                     if (c >= (ndim == 3 ? 8 : 4)) {
                        rs.push_back(sc);
                        continue;
                     }
                     int cc[3] = {gc[0] + off[c][0], gc[1] + off[c][1],
                                  ndim == 3 ? gc[2] + off[c][2] : 0};
                     canon(Hn, l - 1, cc);
                     long long sv = find_cell(Hn, l - 1, cc);
                     if (sv < 0) {
                        set_error("hier_transfer: a level is not nested in the coarser one");
                        return -1;
                     }
                     rs.push_back(sv);
                  }
               }
      }
      DevVec<long long> dcd, dcs, drd, drs;
      DevVec<double> drw;
      auto release = [&]() {
         dcd.release();
         dcs.release();
         drd.release();
         drs.release();
         drw.release();
      };
      if (dcd.upload(cd) || dcs.upload(cs) || drd.upload(rd) || drs.upload(rs) || drw.upload(rw)) {
         release();
         return -1;
      }
      if (dcd.n > 0)
         hipLaunchKernelGGL(k_copy_cells, grid1(dcd.n), dim3(256), 0, s, fo, Ho.ncg, fn, Hn.ncg,
                            ncomp, dcd.p, dcs.p, dcd.n);
      if (drd.n > 0)
         hipLaunchKernelGGL(k_cf_fill<double>, grid1(drd.n), dim3(256), 0, s, fn, Hn.ncg, ncomp, drd.p,
                            drs.p, drw.p, drd.n, 2, ndim);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      release();
      if (e != hipSuccess) {
         set_error(std::string("hier_transfer: ") + hipGetErrorString(e));
         return -1;
      }
   }
   return 0;
}

// New level boxes from the gradient detector, coarsest first (the role of
// GriddingAlgorithm::regridAllFinerLevels in PFModel::Regrid, source/
// PFModel.cc:565-590, with the tags of QuatModel::applyGradientDetector,
// source/QuatModel.cc:3437-3602).  Level 0 is kept.  For k = 0 .. max_levels-2
// the tag field is carried onto the levels 0..k built so far (hier_transfer),
// its ghosts filled by linear refinement, level k tagged, and the tags
// clustered (cluster_mask): over the level-0 domain for k = 0; above, per
// patch of level k inside the patch shrunk by one cell on every side that is
// not a non-periodic physical boundary, so that every new box is properly
// nested.  Clustered level-k boxes refined by 2 are level k+1; no tags end
// the hierarchy.
int hier_regrid(const Hier& Ho, const double* phi_o, const ampe_model_params& p0, double tol,
                const ampe_cluster_params& cp, int max_levels, Hier** out, hipStream_t s)
{
   if (max_levels < 1 || max_levels > kMaxLev) {
      set_error("hier_regrid: max_levels must be 1.." + std::to_string(kMaxLev));
      return -1;
   }
   const int ndim = Ho.ndim;
   std::vector<std::vector<ampe_box>> lev(1);
   lev[0].assign(Ho.boxes.begin() + Ho.pbeg[0], Ho.boxes.begin() + Ho.pbeg[1]);
   int dom0[3] = {Ho.dom[0][0], Ho.dom[0][1], Ho.dom[0][2]};
   double dx0[3] = {Ho.dx[0][0], Ho.dx[0][1], Ho.dx[0][2]};
   auto build = [&](int nl, Hier** H) {
      std::vector<int> np;
      std::vector<ampe_box> flat;
      for (int l = 0; l < nl; l++) {
         np.push_back((int)lev[l].size());
         flat.insert(flat.end(), lev[l].begin(), lev[l].end());
      }
      return hier_create(ndim, nl, np.data(), flat.data(), dom0, dx0, Ho.periodic, H);
   };
   for (int k = 0; k + 1 < max_levels; k++) {
      Hier* Hk = nullptr;
      if (build(k + 1, &Hk)) return -1;
      double* phi = nullptr;
      int* tags = nullptr;
      std::vector<int> ht((size_t)Hk->ncell);
      int rc = 0;
      if (hipMalloc((void**)&phi, sizeof(double) * (size_t)Hk->ncg) != hipSuccess ||
          hipMalloc((void**)&tags, sizeof(int) * (size_t)Hk->ncell) != hipSuccess) {
         set_error("hier_regrid: out of device memory");
         rc = -1;
      }
      if (!rc && fill_zero<double>(phi, Hk->ncg, s)) rc = -1;
      if (!rc && hier_transfer(Ho, *Hk, phi_o, phi, 1, s)) rc = -1;
      if (!rc && hier_ghost_fill(*Hk, phi, 1, 2, s)) rc = -1;
      if (!rc && launch_tag(*Hk->lev[k], lev_params(*Hk, p0, k), phi, tol, tags, s)) rc = -1;
      if (!rc && (hipMemcpyAsync(ht.data(), tags, sizeof(int) * ht.size(), hipMemcpyDeviceToHost,
                                 s) != hipSuccess ||
                  hipStreamSynchronize(s) != hipSuccess)) {
         set_error("hier_regrid: tag copy failed");
         rc = -1;
      }
      (void)hipFree(phi);
      (void)hipFree(tags);
      std::vector<ampe_box> nb;
      for (int p = Hk->pbeg[k]; !rc && p < Hk->pbeg[k + 1]; p++) {
         // the region clustered: the level-0 domain, or the shrunk patch
         ampe_box r;
         if (k == 0) {
            if (p > Hk->pbeg[0]) break;
            for (int a = 0; a < 3; a++) {
               r.lo[a] = 0;
               r.hi[a] = a < ndim ? Hk->dom[0][a] - 1 : 0;
            }
         } else {
            r = Hk->boxes[p];
            bool empty = false;
            for (int a = 0; a < ndim; a++) {
               if (Ho.periodic || r.lo[a] > 0) r.lo[a] += 1;
               if (Ho.periodic || r.hi[a] < Hk->dom[k][a] - 1) r.hi[a] -= 1;
               if (r.hi[a] < r.lo[a]) empty = true;
            }
            if (empty) continue;
         }
         int rn[3] = {box_n(r, 0, ndim), box_n(r, 1, ndim), box_n(r, 2, ndim)};
         std::vector<unsigned char> mask((size_t)rn[0] * rn[1] * rn[2], 0);
         for (int q = Hk->pbeg[k]; q < Hk->pbeg[k + 1]; q++) {
            if (k > 0 && q != p) continue;
            const PatchDev& pd = Hk->lev[k]->hp[q - Hk->pbeg[k]];
            for (int kk = 0; kk < pd.nz; kk++)
               for (int j = 0; j < pd.ny; j++)
                  for (int i = 0; i < pd.nx; i++) {
                     int g[3] = {pd.lo[0] + i, pd.lo[1] + j, pd.lo[2] + kk};
                     bool in = true;
                     for (int a = 0; a < ndim; a++)
                        if (g[a] < r.lo[a] || g[a] > r.hi[a]) in = false;
                     if (!in) continue;
                     long long c =
                         pd.int_off + i + (long long)pd.nx * (j + (long long)pd.ny * kk);
                     if (!ht[(size_t)c]) continue;
                     size_t m = (size_t)(g[0] - r.lo[0]) +
                                (size_t)rn[0] * ((size_t)(g[1] - r.lo[1]) +
                                                 (size_t)rn[1] * (size_t)(ndim == 3 ? g[2] - r.lo[2] : 0));
                     mask[m] = 1;
                  }
         }
         std::vector<ampe_box> bx(mask.size());
         int nbx = 0;
         if (cluster_mask(ndim, r, mask.data(), cp, bx.data(), (int)bx.size(), &nbx)) {
            rc = -1;
            break;
         }
         nb.insert(nb.end(), bx.begin(), bx.begin() + nbx);
      }
      hier_destroy(Hk);
      if (rc) return -1;
      if (nb.empty()) break;
      for (ampe_box& b : nb)
         for (int a = 0; a < ndim; a++) {
            b.lo[a] = 2 * b.lo[a];
            b.hi[a] = 2 * b.hi[a] + 1;
         }
      lev.push_back(nb);
   }
   return build((int)lev.size(), out);
}

}  // namespace ampe

// ------------------------------------------------------------ C ABI
namespace ampe {
ampe_level* level_view(Level* L);  // capi.hip
void level_view_free(ampe_level* v);
}  // namespace ampe

struct ampe_hier {
   ampe::Hier* H;
   ampe_level* views[ampe::kMaxLev + 1];
};

#define AMPE_HREQ(cond, msg)           \
   do {                                 \
      if (!(cond)) {                    \
         ampe::set_error(msg);          \
         return -1;                     \
      }                                 \
   } while (0)

static hipStream_t hstream(ampe_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

namespace ampe {
Hier* hier_of(ampe_hier* h) { return h ? h->H : nullptr; }
}  // namespace ampe

extern "C" {

static ampe_hier* hier_wrap(ampe::Hier* H)
{
   ampe_hier* h = new ampe_hier();
   h->H = H;
   for (int l = 0; l <= ampe::kMaxLev; l++) h->views[l] = nullptr;
   for (int l = 0; l < H->nlev; l++) h->views[l] = ampe::level_view(H->lev[l]);
   h->views[ampe::kMaxLev] = ampe::level_view(H->comp);
   return h;
}

int ampe_hier_create(int ndim, int nlev, const int* npatch, const ampe_box* boxes,
                     const int dom0[3], const double dx0[3], int periodic, ampe_hier** out)
{
   AMPE_HREQ(out, "null argument");
   ampe::Hier* H = nullptr;
   if (ampe::hier_create(ndim, nlev, npatch, boxes, dom0, dx0, periodic, &H)) return -1;
   *out = hier_wrap(H);
   return 0;
}

int ampe_hier_boxes(const ampe_hier* h, int* nlev, int npatch[AMPE_MAX_LEVELS], ampe_box* boxes,
                    int max_boxes)
{
   AMPE_HREQ(h && nlev && npatch, "null argument");
   const ampe::Hier& H = *h->H;
   *nlev = H.nlev;
   for (int l = 0; l < AMPE_MAX_LEVELS; l++)
      npatch[l] = l < H.nlev ? H.pbeg[l + 1] - H.pbeg[l] : 0;
   AMPE_HREQ((int)H.boxes.size() <= max_boxes || !boxes, "ampe_hier_boxes: max_boxes too small");
   if (boxes)
      for (size_t b = 0; b < H.boxes.size(); b++) boxes[b] = H.boxes[b];
   return 0;
}

int ampe_hier_transfer(const ampe_hier* old_h, const ampe_hier* new_h, const double* f_old,
                       double* f_new, int ncomp, ampe_stream_t stream)
{
   AMPE_HREQ(old_h && new_h && f_old && f_new, "null argument");
   return ampe::hier_transfer(*old_h->H, *new_h->H, f_old, f_new, ncomp, hstream(stream));
}

int ampe_hier_regrid(const ampe_hier* old_h, const double* phi, const ampe_model_params* p0,
                     double tol, const ampe_cluster_params* cp, int max_levels, ampe_hier** out,
                     ampe_stream_t stream)
{
   AMPE_HREQ(old_h && phi && p0 && cp && out, "null argument");
   AMPE_HREQ(cp->efficiency_tol > 0.0 && cp->efficiency_tol <= 1.0 && cp->tag_buffer >= 0,
             "ampe_hier_regrid: need 0 < efficiency_tol <= 1 and tag_buffer >= 0");
   ampe::Hier* H = nullptr;
   if (ampe::hier_regrid(*old_h->H, phi, *p0, tol, *cp, max_levels, &H, hstream(stream)))
      return -1;
   *out = hier_wrap(H);
   return 0;
}

int ampe_hier_destroy(ampe_hier* h)
{
   if (!h) return 0;
   for (int l = 0; l <= ampe::kMaxLev; l++)
      if (h->views[l]) ampe::level_view_free(h->views[l]);
   ampe::hier_destroy(h->H);
   delete h;
   return 0;
}

int ampe_hier_sizes(const ampe_hier* h, long long* ncell_ghost, long long* ncell,
                    long long nside[3])
{
   AMPE_HREQ(h, "null argument");
   if (ncell_ghost) *ncell_ghost = h->H->ncg;
   if (ncell) *ncell = h->H->ncell;
   if (nside)
      for (int a = 0; a < 3; a++) nside[a] = h->H->nside[a];
   return 0;
}

int ampe_hier_level(ampe_hier* h, int l, ampe_level** level)
{
   AMPE_HREQ(h && level && l >= 0 && l < h->H->nlev, "bad level");
   *level = h->views[l];
   return 0;
}

int ampe_hier_composite(ampe_hier* h, ampe_level** level)
{
   AMPE_HREQ(h && level, "null argument");
   *level = h->views[ampe::kMaxLev];
   return 0;
}

int ampe_hier_weights(const ampe_hier* h, double* w, ampe_stream_t stream)
{
   AMPE_HREQ(h && w, "null argument");
   if (hipMemcpyAsync(w, h->H->wt, sizeof(double) * h->H->ncg, hipMemcpyDeviceToDevice,
                      hstream(stream)) != hipSuccess) {
      ampe::set_error("ampe_hier_weights: copy failed");
      return -1;
   }
   return 0;
}

int ampe_hier_ghost_fill(const ampe_hier* h, double* f, int ncomp, int cf, ampe_stream_t stream)
{
   AMPE_HREQ(h && f && ncomp > 0 && cf >= 0 && cf <= 2, "bad argument");
   return ampe::hier_ghost_fill(*h->H, f, ncomp, cf, hstream(stream));
}

int ampe_hier_coarsen(const ampe_hier* h, double* f, int ncomp, int quatnorm, ampe_stream_t stream)
{
   AMPE_HREQ(h && f && ncomp > 0, "bad argument");
   return ampe::hier_coarsen(*h->H, f, ncomp, quatnorm, hstream(stream));
}

int ampe_hier_phase_rhs(const ampe_hier* h, const ampe_model_params* p0, const double* phi,
                        const double* q, const double* cl, const double* cs, double* rhs,
                        ampe_stream_t stream)
{
   AMPE_HREQ(h && p0 && phi && rhs, "null argument");
   return ampe::hier_phase_rhs(*h->H, *p0, phi, q, cl, cs, rhs, hstream(stream));
}

int ampe_hier_quat_face_norm2(const ampe_hier* h, const ampe_model_params* p0, const double* q,
                              double* n2, ampe_stream_t stream)
{
   AMPE_HREQ(h && p0 && q && n2, "null argument");
   return ampe::hier_quat_face_norm2(*h->H, *p0, q, n2, hstream(stream));
}

int ampe_hier_quat_rhs(const ampe_hier* h, const ampe_model_params* p0, const double* phi,
                       const double* q, const double* n2, double* fc, double* rhs,
                       ampe_stream_t stream)
{
   AMPE_HREQ(h && p0 && phi && q && n2 && fc && rhs, "null argument");
   return ampe::hier_quat_rhs(*h->H, *p0, phi, q, n2, fc, rhs, hstream(stream));
}

int ampe_hier_conc_coeffs(const ampe_hier* h, const ampe_model_params* p0, const double* phi,
                          const double* cl, const double* cs, double* Dc, double* Dphi,
                          ampe_stream_t stream)
{
   AMPE_HREQ(h && p0 && phi && cl && cs && Dc && Dphi, "null argument");
   return ampe::hier_conc_coeffs(*h->H, *p0, phi, cl, cs, Dc, Dphi, hstream(stream));
}

int ampe_hier_conc_rhs(const ampe_hier* h, const ampe_model_params* p0, const double* c,
                       const double* phi, const double* Dc, const double* Dphi, double* rhs,
                       ampe_stream_t stream)
{
   AMPE_HREQ(h && p0 && c && phi && Dc && Dphi && rhs, "null argument");
   return ampe::hier_conc_rhs(*h->H, *p0, c, phi, Dc, Dphi, rhs, hstream(stream));
}

int ampe_hier_fac_solve(ampe_hier* h, const ampe_hier_block* b, const ampe_fac_params* fp,
                        const double* f, double* u, int ncomp, double tol, double* norm,
                        int* cycles, ampe_stream_t stream)
{
   AMPE_HREQ(h && b && fp && f && u, "null argument");
   AMPE_HREQ(fp->max_cycles >= 1 && fp->pre >= 0 && fp->post >= 0 && fp->coarse_its >= 0,
             "bad FAC parameters");
   return ampe::hier_fac_solve(*h->H, *b, *fp, f, u, ncomp, tol, norm, cycles, hstream(stream));
}

}  // extern "C"
