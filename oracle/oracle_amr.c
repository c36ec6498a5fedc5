/* Multi-level (AMR) restatement -- TEST INFRASTRUCTURE ONLY.
 *
 * A hierarchy of levels with refinement ratio 2: level 0 covers the domain,
 * level l+1 is a set of boxes (patches) in level-(l+1) index space, aligned
 * with level-l cells and properly nested (every fine patch grown by one
 * coarse cell lies on level l, except across a physical boundary).  Fields
 * are composite arrays: per component, the ghosted blocks (ghost width 1) of
 * every patch of every level, level by level; side fields hold every patch's
 * face block (x faces, y faces, z faces of the patch in a row).
 *
 * What is restated, each function citing the reference:
 *  - ghost fill of a level from the level itself, the periodic images or the
 *    zero-slope boundary, and from the next coarser level on the coarse-fine
 *    boundary by constant or linear refinement (SAMRAI refine schedules
 *    created in QuatIntegrator::fillScratch, source/QuatIntegrator.cc:
 *    2873-2957, with LINEAR_REFINE for phase, composition and orientation,
 *    QuatModel.cc:924-955; the FAC operators' CONSTANT_REFINE ghost fill,
 *    source/EllipticFACOps.cc:612-616);
 *  - conservative coarsening of cells and faces (QuatIntegrator::coarsenData,
 *    :2289-2331; the flux coarsen schedules, :945-955);
 *  - the composite right-hand sides: finest to coarsest, fluxes of a level,
 *    fine fluxes coarsened onto the coarse faces, then the divergence
 *    (PhaseRHSStrategyWithQ::evaluateRHS, source/PhaseRHSStrategyWithQ.cc:
 *    90-149; QuatIntegrator::evaluateConcentrationRHS, :2578-2690;
 *    QuatFACOps::evaluateRHS -> accumulateOperatorOnLevel, source/
 *    QuatFACOps.cc:1861-1890);
 *  - the FAC preconditioner over levels: FACPreconditioner::solveSystem and
 *    facCycle_Recursive (source/samrai/FACPreconditioner.cc:267-474,
 *    661-680) with EllipticFACOps' operations (source/EllipticFACOps.cc:
 *    948-1230 smoothing, restriction, prolongation; :1239 Ewing's
 *    coarse-fine flux fix, F77 efo_ewingfixflux*, source/fortran/2d/
 *    ellipticfacops.m4:543-780, 3d/ellipticfacops.m4:906-1200; :1480-1760
 *    composite residual and RMS norm).
 *
 * The reference's FP64 arithmetic; fluxes are materialised in side arrays as
 * the reference does.  Only cells not covered by a finer level carry
 * information (their weights are zero, and coarsenData overwrites them):
 * the device restatement and this one are compared there.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OH_MAXLEV 4

struct OHier {
   int ndim, nlev, periodic, np;
   int dom[OH_MAXLEV][3];       /* level domain extents (cells) */
   double dx[OH_MAXLEV][3];
   int pbeg[OH_MAXLEV + 1];     /* patches of level l: [pbeg[l], pbeg[l+1]) */
   OGeom* G;                    /* per patch, with its level's dx */
   int (*lo)[3];                /* per patch lower corner (level index space) */
   long* coff;                  /* per patch ghosted-block offset in a component block */
   long* foff;                  /* per patch face-block offset in a side field */
   long ncg, nf;
   /* ghost tables, per level: entries [gbeg[l], gbeg[l+1]) */
   long* gbeg;
   long* gdst;                  /* ghost cell (composite index) */
   int* gkind;                  /* 0 same level, 1 coarse-fine, -1 none */
   long* gsrc;                  /* same level: [0]; coarse-fine: parent [0], linear corners [1..8] */
   double* gw;                  /* linear weights x, y, z */
   long* gfine;                 /* coarse-fine: the clamped fine cell's global index (for Ewing) */
};

static int lev_of(const OHier* H, int p)
{
   int l = 0;
   while (p >= H->pbeg[l + 1]) l++;
   return l;
}

static long cidx(const OHier* H, int p, int i, int j, int k)
{
   const OGeom* G = &H->G[p];
   return H->coff[p] + (long)(i + G->g) + (long)(j + G->g) * G->sy + (long)(k + G->gz) * G->sz;
}

static void canon(const OHier* H, int l, int g[3])
{
   for (int a = 0; a < H->ndim; a++) {
      int n = H->dom[l][a];
      if (H->periodic) {
         g[a] = ((g[a] % n) + n) % n;
      } else {
         if (g[a] < 0) g[a] = 0;
         if (g[a] > n - 1) g[a] = n - 1;
      }
   }
   if (H->ndim == 2) g[2] = 0;
}

static int inbox(const OHier* H, int p, const int g[3])
{
   const OGeom* G = &H->G[p];
   int n[3] = {G->nx, G->ny, G->nz};
   for (int a = 0; a < 3; a++)
      if (g[a] < H->lo[p][a] || g[a] >= H->lo[p][a] + n[a]) return 0;
   return 1;
}

/* composite index of level-l cell g (canonical), -1 if no patch of l holds it */
static long find(const OHier* H, int l, const int g[3])
{
   for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++)
      if (inbox(H, p, g))
         return cidx(H, p, g[0] - H->lo[p][0], g[1] - H->lo[p][1], g[2] - H->lo[p][2]);
   return -1;
}

/* quatlinrefine's coarse index and weight of a fine index (ratio 2),
 * restated as in the single-level linear refine */
static void lin_index(int gf, double dxf, double dxc, int* gc, double* x)
{
   int ic = gf < 0 ? (gf + 1) / 2 - 1 : gf / 2;
   int ir = gf - ic * 2;
   double dlt = ((double)ir + 0.5) * dxf - dxc * 0.5;
   double xx = dlt / dxc;
   if (xx < 0.0) {
      ic = ic - 1;
      xx = xx + 1.0;
   }
   *gc = ic;
   *x = xx;
}

static int fdiv2(int a) { return a < 0 ? (a + 1) / 2 - 1 : a / 2; }

void o_hier_destroy(OHier* H)
{
   if (!H) return;
   free(H->G);
   free(H->lo);
   free(H->coff);
   free(H->foff);
   free(H->gbeg);
   free(H->gdst);
   free(H->gkind);
   free(H->gsrc);
   free(H->gw);
   free(H->gfine);
   free(H);
}

OHier* o_hier_create(int ndim, int nlev, const int* npatch, const ampe_box* boxes,
                     const int dom0[3], const double dx0[3], int periodic)
{
   if (nlev < 1 || nlev > OH_MAXLEV || (ndim != 2 && ndim != 3)) return NULL;
   OHier* H = (OHier*)calloc(1, sizeof(OHier));
   H->ndim = ndim;
   H->nlev = nlev;
   H->periodic = periodic ? 1 : 0;
   int np = 0;
   for (int l = 0; l < nlev; l++) {
      H->pbeg[l] = np;
      np += npatch[l];
      for (int a = 0; a < 3; a++) {
         H->dom[l][a] = (a < ndim) ? dom0[a] << l : 1;
         H->dx[l][a] = (a < ndim) ? dx0[a] / (double)(1 << l) : 1.0;
      }
   }
   H->pbeg[nlev] = np;
   H->np = np;
   H->G = (OGeom*)calloc(np, sizeof(OGeom));
   H->lo = calloc(np, sizeof(int[3]));
   H->coff = (long*)calloc(np, sizeof(long));
   H->foff = (long*)calloc(np, sizeof(long));
   long co = 0, fo = 0;
   for (int p = 0; p < np; p++) {
      int l = lev_of(H, p);
      int n[3];
      for (int a = 0; a < 3; a++) {
         H->lo[p][a] = (a < ndim) ? boxes[p].lo[a] : 0;
         n[a] = (a < ndim) ? boxes[p].hi[a] - boxes[p].lo[a] + 1 : 1;
         if (n[a] < 1 || H->lo[p][a] < 0 || H->lo[p][a] + n[a] > H->dom[l][a]) {
            o_hier_destroy(H);
            return NULL;
         }
         /* aligned with the coarser level's cells */
         if (l > 0 && a < ndim && ((H->lo[p][a] & 1) || (n[a] & 1))) {
            o_hier_destroy(H);
            return NULL;
         }
      }
      o_geom_init(&H->G[p], ndim, n[0], n[1], n[2], 1, H->dx[l]);
      H->G[p].periodic = 0;
      H->coff[p] = co;
      H->foff[p] = fo;
      co += H->G[p].ncg;
      fo += H->G[p].nfx + H->G[p].nfy + H->G[p].nfz;
   }
   H->ncg = co;
   H->nf = fo;
   /* level 0 covers the domain; same-level patches are disjoint */
   for (int l = 0; l < nlev; l++) {
      long cells = 0;
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         cells += (long)H->G[p].nx * H->G[p].ny * H->G[p].nz;
         for (int q = H->pbeg[l]; q < p; q++) {
            int ov = 1;
            for (int a = 0; a < 3; a++) {
               int na[3] = {H->G[p].nx, H->G[p].ny, H->G[p].nz};
               int nb[3] = {H->G[q].nx, H->G[q].ny, H->G[q].nz};
               if (H->lo[p][a] + na[a] <= H->lo[q][a] || H->lo[q][a] + nb[a] <= H->lo[p][a])
                  ov = 0;
            }
            if (ov) {
               o_hier_destroy(H);
               return NULL;
            }
         }
      }
      if (l == 0 && cells != (long)H->dom[0][0] * H->dom[0][1] * H->dom[0][2]) {
         o_hier_destroy(H);
         return NULL;
      }
   }
   /* ghost tables */
   long cap = 0;
   for (int p = 0; p < np; p++) {
      const OGeom* G = &H->G[p];
      cap += G->ncg - (long)G->nx * G->ny * G->nz;
   }
   H->gbeg = (long*)calloc(nlev + 1, sizeof(long));
   H->gdst = (long*)calloc(cap + 1, sizeof(long));
   H->gkind = (int*)calloc(cap + 1, sizeof(int));
   H->gsrc = (long*)calloc(9 * (cap + 1), sizeof(long));
   H->gw = (double*)calloc(3 * (cap + 1), sizeof(double));
   H->gfine = (long*)calloc(cap + 1, sizeof(long));
   long ne = 0;
   for (int l = 0; l < nlev; l++) {
      H->gbeg[l] = ne;
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         const OGeom* G = &H->G[p];
         for (int k = -G->gz; k < G->nz + G->gz; k++)
            for (int j = -1; j < G->ny + 1; j++)
               for (int i = -1; i < G->nx + 1; i++) {
                  if (i >= 0 && i < G->nx && j >= 0 && j < G->ny && k >= 0 && k < G->nz)
                     continue;
                  int g[3] = {H->lo[p][0] + i, H->lo[p][1] + j, H->lo[p][2] + k};
                  canon(H, l, g);
                  long e = ne++;
                  H->gdst[e] = cidx(H, p, i, j, k);
                  H->gfine[e] = ((long)g[2] * H->dom[l][1] + g[1]) * H->dom[l][0] + g[0];
                  long s = find(H, l, g);
                  if (s >= 0) {
                     H->gkind[e] = 0;
                     H->gsrc[9 * e] = s;
                     continue;
                  }
                  if (l == 0) {
                     o_hier_destroy(H);
                     return NULL;
                  }
                  H->gkind[e] = 1;
                  int cp[3] = {fdiv2(g[0]), fdiv2(g[1]), ndim == 3 ? fdiv2(g[2]) : 0};
                  canon(H, l - 1, cp);
                  long sc = find(H, l - 1, cp);
                  if (sc < 0) {
                     o_hier_destroy(H);
                     return NULL;
                  }
                  H->gsrc[9 * e] = sc;
                  int gc[3] = {0, 0, 0};
                  double w[3] = {0.0, 0.0, 0.0};
                  for (int a = 0; a < ndim; a++)
                     lin_index(g[a], H->dx[l][a], H->dx[l - 1][a], &gc[a], &w[a]);
                  for (int a = 0; a < 3; a++) H->gw[3 * e + a] = w[a];
                  /* corners q1..q8 in quatlinrefine's order: (0,0,0), (0,1,0),
                   * (1,0,0), (1,1,0), then the same at +1 in z */
                  const int off[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0},
                                         {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
                  int nc = ndim == 3 ? 8 : 4;
                  for (int c = 0; c < nc; c++) {
                     int cc[3] = {gc[0] + off[c][0], gc[1] + off[c][1],
                                  ndim == 3 ? gc[2] + off[c][2] : 0};
                     canon(H, l - 1, cc);
                     long sv = find(H, l - 1, cc);
                     if (sv < 0) {
                        o_hier_destroy(H);
                        return NULL;
                     }
                     H->gsrc[9 * e + 1 + c] = sv;
                  }
               }
      }
   }
   H->gbeg[nlev] = ne;
   return H;
}

/* The regrid fill of QuatModel::Regrid (source/QuatModel.cc:2733-2790; the
 * refine schedule of initializeLevelData, :2997-3004): a cell of the new
 * hierarchy's level l takes the old level l's value where an old patch holds
 * it, else LINEAR_REFINE of the new level l-1 (quatlinrefine's interpolation
 * of the coarse-fine ghost fill), coarsest level first.  Interior cells only;
 * returns -1 when the hierarchies differ in dimension, level-0 domain or
 * periodicity, or a cell finds no source. */
int o_hier_transfer(const OHier* Ho, const OHier* Hn, const double* fo, double* fn, int ncomp)
{
   int nd = Hn->ndim;
   if (Ho->ndim != nd || Ho->periodic != Hn->periodic) return -1;
   for (int a = 0; a < 3; a++)
      if (Ho->dom[0][a] != Hn->dom[0][a] || Ho->dx[0][a] != Hn->dx[0][a]) return -1;
   const int off[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0},
                          {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
   for (int l = 0; l < Hn->nlev; l++)
      for (int p = Hn->pbeg[l]; p < Hn->pbeg[l + 1]; p++) {
         const OGeom* G = &Hn->G[p];
         for (int k = 0; k < G->nz; k++)
            for (int j = 0; j < G->ny; j++)
               for (int i = 0; i < G->nx; i++) {
                  int g[3] = {Hn->lo[p][0] + i, Hn->lo[p][1] + j, Hn->lo[p][2] + k};
                  long dst = cidx(Hn, p, i, j, k);
                  long src = l < Ho->nlev ? find(Ho, l, g) : -1;
                  if (src >= 0) {
                     for (int m = 0; m < ncomp; m++)
                        fn[(long)m * Hn->ncg + dst] = fo[(long)m * Ho->ncg + src];
                     continue;
                  }
                  if (l == 0) return -1;
                  int gc[3] = {0, 0, 0};
                  double w[3] = {0.0, 0.0, 0.0};
                  for (int a = 0; a < nd; a++)
                     lin_index(g[a], Hn->dx[l][a], Hn->dx[l - 1][a], &gc[a], &w[a]);
                  long sv[8];
                  int nc = nd == 3 ? 8 : 4;
                  for (int c = 0; c < nc; c++) {
                     int cc[3] = {gc[0] + off[c][0], gc[1] + off[c][1],
                                  nd == 3 ? gc[2] + off[c][2] : 0};
                     canon(Hn, l - 1, cc);
                     sv[c] = find(Hn, l - 1, cc);
                     if (sv[c] < 0) return -1;
                  }
                  for (int m = 0; m < ncomp; m++) {
                     const double* a = fn + (long)m * Hn->ncg;
                     double x = w[0], y = w[1], z = w[2];
                     double q1 = a[sv[0]], q2 = a[sv[1]], q3 = a[sv[2]], q4 = a[sv[3]];
                     double v;
                     if (nd == 2) {
                        v = (q1 + (q3 - q1) * x) * (1.0 - y) + (q2 + (q4 - q2) * x) * y;
                     } else {
                        double q5 = a[sv[4]], q6 = a[sv[5]], q7 = a[sv[6]], q8 = a[sv[7]];
                        v = ((q1 + (q3 - q1) * x) * (1.0 - y) + (q2 + (q4 - q2) * x) * y) *
                                (1.0 - z) +
                            ((q5 + (q7 - q5) * x) * (1.0 - y) + (q6 + (q8 - q6) * x) * y) * z;
                     }
                     fn[(long)m * Hn->ncg + dst] = v;
                  }
               }
      }
   return 0;
}

int o_hier_nlev(const OHier* H) { return H->nlev; }
int o_hier_npatch(const OHier* H) { return H->np; }
long o_hier_ncg(const OHier* H) { return H->ncg; }
long o_hier_nf(const OHier* H) { return H->nf; }
void o_hier_offsets(const OHier* H, long* coff, long* foff)
{
   for (int p = 0; p < H->np; p++) {
      if (coff) coff[p] = H->coff[p];
      if (foff) foff[p] = H->foff[p];
   }
}
const OGeom* o_hier_geom(const OHier* H, int p) { return &H->G[p]; }

/* Ghost cells of level l: from the level (periodic image or zero-slope
 * image), and on the coarse-fine boundary (cf = 0: left as they are, the
 * "NoCoarse" fill; 1: CONSTANT_REFINE, the coarse parent's value; 2:
 * LINEAR_REFINE, quatlinrefine's interpolation of the coarse cells). */
void o_hier_ghost_fill_level(const OHier* H, int l, double* f, int ncomp, int cf)
{
   for (int m = 0; m < ncomp; m++) {
      double* a = f + (long)m * H->ncg;
      for (long e = H->gbeg[l]; e < H->gbeg[l + 1]; e++) {
         const long* s = H->gsrc + 9 * e;
         if (H->gkind[e] == 0) {
            a[H->gdst[e]] = a[s[0]];
         } else if (cf == 1) {
            a[H->gdst[e]] = a[s[0]];
         } else if (cf == 2) {
            double x = H->gw[3 * e], y = H->gw[3 * e + 1], z = H->gw[3 * e + 2];
            double q1 = a[s[1]], q2 = a[s[2]], q3 = a[s[3]], q4 = a[s[4]];
            double v;
            if (H->ndim == 2) {
               v = (q1 + (q3 - q1) * x) * (1.0 - y) + (q2 + (q4 - q2) * x) * y;
            } else {
               double q5 = a[s[5]], q6 = a[s[6]], q7 = a[s[7]], q8 = a[s[8]];
               v = ((q1 + (q3 - q1) * x) * (1.0 - y) + (q2 + (q4 - q2) * x) * y) * (1.0 - z) +
                   ((q5 + (q7 - q5) * x) * (1.0 - y) + (q6 + (q8 - q6) * x) * y) * z;
            }
            a[H->gdst[e]] = v;
         }
      }
   }
}

/* every level, coarsest first (the coarse ghosts a linear refine reads are
 * coarse interior cells, by the nesting) */
void o_hier_ghost_fill(const OHier* H, double* f, int ncomp, int cf)
{
   for (int l = 0; l < H->nlev; l++) o_hier_ghost_fill_level(H, l, f, ncomp, cf);
}

/* covered[n] = 1 for interior cells of level l lying under level l+1 */
static void covered_mask(const OHier* H, int l, unsigned char* cov)
{
   memset(cov, 0, (size_t)H->ncg);
   if (l + 1 >= H->nlev) return;
   for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
      const OGeom* G = &H->G[p];
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               int g[3] = {2 * (H->lo[p][0] + i), 2 * (H->lo[p][1] + j),
                           H->ndim == 3 ? 2 * (H->lo[p][2] + k) : 0};
               if (find(H, l + 1, g) >= 0) cov[cidx(H, p, i, j, k)] = 1;
            }
   }
}

/* Control-volume weights of interior cells: the cell volume relative to a
 * level-0 cell, zero under a finer level (the weight field of
 * QuatModel::computeVectorWeights, source/QuatModel.cc:4710-4750); ghost
 * entries zero.  On one level every weight is 1. */
void o_hier_weights(const OHier* H, double* w)
{
   memset(w, 0, sizeof(double) * H->ncg);
   unsigned char* cov = (unsigned char*)malloc((size_t)H->ncg);
   for (int l = 0; l < H->nlev; l++) {
      covered_mask(H, l, cov);
      double v = 1.0;
      for (int a = 0; a < H->ndim; a++) v = v * 0.5;
      double cv = 1.0;
      for (int r = 0; r < l; r++) cv = cv * v;
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         const OGeom* G = &H->G[p];
         for (int k = 0; k < G->nz; k++)
            for (int j = 0; j < G->ny; j++)
               for (int i = 0; i < G->nx; i++) {
                  long c = cidx(H, p, i, j, k);
                  w[c] = cov[c] ? 0.0 : cv;
               }
      }
   }
   free(cov);
}

/* CONSERVATIVE_COARSEN of cells from level l+1 onto the level-l cells it
 * covers: sum of child * dV_fine, divided by dV_coarse.  quatnorm: then
 * q / |q| over the ncomp components (QuatWeightedAverage). */
void o_hier_coarsen_level(const OHier* H, int l, double* f, int ncomp, int quatnorm)
{
   if (l + 1 >= H->nlev) return;
   double dvf = H->dx[l + 1][0] * H->dx[l + 1][1] * (H->ndim == 3 ? H->dx[l + 1][2] : 1.0);
   double dvc = H->dx[l][0] * H->dx[l][1] * (H->ndim == 3 ? H->dx[l][2] : 1.0);
   int nz = H->ndim == 3 ? 2 : 1;
   for (int p = H->pbeg[l + 1]; p < H->pbeg[l + 2]; p++) {
      const OGeom* G = &H->G[p];
      for (int ck = 0; ck < (H->ndim == 3 ? G->nz / 2 : 1); ck++)
         for (int cj = 0; cj < G->ny / 2; cj++)
            for (int ci = 0; ci < G->nx / 2; ci++) {
               int g[3] = {H->lo[p][0] / 2 + ci, H->lo[p][1] / 2 + cj,
                           H->ndim == 3 ? H->lo[p][2] / 2 + ck : 0};
               long c0 = find(H, l, g);
               if (c0 < 0) continue;
               double acc[8];
               for (int m = 0; m < ncomp && m < 8; m++) {
                  const double* a = f + (long)m * H->ncg;
                  double s = 0.0;
                  for (int r2 = 0; r2 < nz; r2++)
                     for (int r1 = 0; r1 < 2; r1++)
                        for (int r0 = 0; r0 < 2; r0++)
                           s = s + a[cidx(H, p, 2 * ci + r0, 2 * cj + r1,
                                          H->ndim == 3 ? 2 * ck + r2 : 0)] * dvf;
                  acc[m] = s / dvc;
               }
               if (quatnorm) {
                  double s = 0.0;
                  for (int m = 0; m < ncomp; m++) s = s + acc[m] * acc[m];
                  double mg = sqrt(s);
                  double minv = mg < 1.0e-15 ? 0.0 : 1.0 / mg;
                  for (int m = 0; m < ncomp; m++) acc[m] = acc[m] * minv;
               }
               for (int m = 0; m < ncomp && m < 8; m++) f[(long)m * H->ncg + c0] = acc[m];
            }
   }
}

/* QuatIntegrator::coarsenData: finest to coarsest */
void o_hier_coarsen(const OHier* H, double* f, int ncomp, int quatnorm)
{
   for (int l = H->nlev - 2; l >= 0; l--) o_hier_coarsen_level(H, l, f, ncomp, quatnorm);
}

/* face index of patch p, direction d, face (i, j, k) (face i along x etc.) */
static long fidx(const OHier* H, int p, int d, int i, int j, int k)
{
   const OGeom* G = &H->G[p];
   if (d == 0) return H->foff[p] + (long)i + (long)(G->nx + 1) * ((long)j + (long)G->ny * k);
   if (d == 1)
      return H->foff[p] + G->nfx + (long)i + (long)G->nx * ((long)j + (long)(G->ny + 1) * k);
   return H->foff[p] + G->nfx + G->nfy + (long)i + (long)G->nx * ((long)j + (long)G->ny * k);
}

/* Side-data CONSERVATIVE_COARSEN from the faces of level l+1 onto the
 * level-l faces they cover (every face of a fine patch's coarsened box):
 * sum of fine face * fine face area, divided by the coarse face area.  A
 * coarse face shared by two coarse patches is written in both. */
/* coarse faces under the level-(l+1) patches take the area-weighted average
 * of their fine faces; outer != 0 restricts this to the coarse faces on the
 * fine patch boundaries (an OutersideData source: the FAC composite residual,
 * EllipticFACOps.cc:1524-1528), outer == 0 takes every face under a fine patch
 * (a SideData source: the right-hand sides' flux coarsening) */
static void coarsen_faces(const OHier* H, int l, const double* src, double* dst, int ncomp,
                          long stride, int outer)
{
   if (l + 1 >= H->nlev) return;
   int nd = H->ndim;
   for (int p = H->pbeg[l + 1]; p < H->pbeg[l + 2]; p++) {
      const OGeom* G = &H->G[p];
      int cn[3] = {G->nx / 2, G->ny / 2, nd == 3 ? G->nz / 2 : 1};
      int clo[3] = {H->lo[p][0] / 2, H->lo[p][1] / 2, nd == 3 ? H->lo[p][2] / 2 : 0};
      for (int d = 0; d < nd; d++) {
         int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
         if (nd == 2) {
            t1 = 1 - d;
            t2 = 2;
         }
         double af = 1.0, ac = 1.0;
         for (int a = 0; a < nd; a++)
            if (a != d) {
               af = af * H->dx[l + 1][a];
               ac = ac * H->dx[l][a];
            }
         int e[3] = {cn[0] + (d == 0), cn[1] + (d == 1), cn[2] + (d == 2)};
         for (int ck = 0; ck < e[2]; ck++)
            for (int cj = 0; cj < e[1]; cj++)
               for (int ci = 0; ci < e[0]; ci++) {
                  int cf[3] = {ci, cj, ck};
                  if (outer && cf[d] != 0 && cf[d] != cn[d]) continue;
                  int gcf[3] = {clo[0] + ci, clo[1] + cj, clo[2] + ck};
                  for (int m = 0; m < ncomp; m++) {
                     const double* a = src + (long)m * stride;
                     double* b = dst + (long)m * stride;
                     double acc = 0.0;
                     int n2 = (nd == 3) ? 2 : 1;
                     for (int r2 = 0; r2 < n2; r2++)
                        for (int r1 = 0; r1 < 2; r1++) {
                           int ff[3];
                           ff[d] = 2 * cf[d];
                           ff[t1] = 2 * cf[t1] + r1;
                           if (nd == 3) ff[t2] = 2 * cf[t2] + r2;
                           else ff[2] = 0;
                           acc = acc + a[fidx(H, p, d, ff[0], ff[1], ff[2])] * af;
                        }
                     double v = acc / ac;
                     /* every coarse patch holding this face */
                     for (int q = H->pbeg[l]; q < H->pbeg[l + 1]; q++) {
                        const OGeom* C = &H->G[q];
                        int qn[3] = {C->nx, C->ny, C->nz};
                        int loc[3];
                        int ok = 1;
                        for (int b = 0; b < 3; b++) {
                           loc[b] = gcf[b] - H->lo[q][b];
                           int ext = qn[b] + (b == d);
                           if (loc[b] < 0 || loc[b] >= ext) ok = 0;
                        }
                        if (ok) b[fidx(H, q, d, loc[0], loc[1], loc[2])] = v;
                     }
                  }
               }
      }
   }
}

void o_hier_coarsen_faces_to(const OHier* H, int l, const double* src, double* dst, int ncomp,
                             long stride)
{
   coarsen_faces(H, l, src, dst, ncomp, stride, 0);
}

void o_hier_coarsen_faces(const OHier* H, int l, double* s, int ncomp, long stride)
{
   coarsen_faces(H, l, s, s, ncomp, stride, 0);
}

/* per-level copy of the model parameters with the level's mesh spacing */
static ampe_model_params lev_params(const OHier* H, const ampe_model_params* p0, int l)
{
   ampe_model_params p = *p0;
   for (int a = 0; a < 3; a++) p.dx[a] = H->dx[l][a];
   return p;
}

/* ---------------------------------------------------------------- RHS */

/* patch p's blocks of an ncomp-component composite field, stacked at the
 * patch's own component stride (the layout of the per-patch kernels) */
static double* pgather(const OHier* H, int p, const double* f, int ncomp)
{
   const OGeom* G = &H->G[p];
   double* b = (double*)malloc(sizeof(double) * G->ncg * (ncomp > 0 ? ncomp : 1));
   for (int m = 0; m < ncomp; m++)
      memcpy(b + (long)m * G->ncg, f + (long)m * H->ncg + H->coff[p], sizeof(double) * G->ncg);
   return b;
}

static void pscatter(const OHier* H, int p, const double* b, double* f, int ncomp)
{
   const OGeom* G = &H->G[p];
   for (int m = 0; m < ncomp; m++) {
      const double* src = b + (long)m * G->ncg;
      double* dst = f + (long)m * H->ncg + H->coff[p];
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++)
            for (int i = 0; i < G->nx; i++) {
               long c = (long)(i + G->g) + (long)(j + G->g) * G->sy + (long)(k + G->gz) * G->sz;
               dst[c] = src[c];
            }
   }
}

static long pnf(const OGeom* G) { return G->nfx + G->nfy + G->nfz; }

/* faces of patch p of an ncomp-component composite side field (stride H->nf) */
static double* fgather(const OHier* H, int p, const double* s, int ncomp)
{
   long n = pnf(&H->G[p]);
   double* b = (double*)malloc(sizeof(double) * n * (ncomp > 0 ? ncomp : 1));
   for (int m = 0; m < ncomp; m++)
      memcpy(b + (long)m * n, s + (long)m * H->nf + H->foff[p], sizeof(double) * n);
   return b;
}

static void fscatter(const OHier* H, int p, const double* b, double* s, int ncomp)
{
   long n = pnf(&H->G[p]);
   for (int m = 0; m < ncomp; m++)
      memcpy(s + (long)m * H->nf + H->foff[p], b + (long)m * n, sizeof(double) * n);
}

/* Composite phase RHS (ghosts of phi and q filled on every level): finest to
 * coarsest, the level's side fluxes, the finer level's fluxes coarsened onto
 * it, then computerhspbg per patch (source/PhaseRHSStrategyWithQ.cc:126-144). */
void o_hier_phase_rhs(const OHier* H, const ampe_model_params* p0, const double* phi,
                      const double* q, const double* cl, const double* cs, double* rhs)
{
   double* flux = (double*)calloc((size_t)H->nf, sizeof(double));
   int ql = p0->qlen;
   for (int l = H->nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* qb = q ? pgather(H, pp, q, ql) : NULL;
         double* fb = (double*)malloc(sizeof(double) * pnf(&H->G[pp]));
         o_phase_flux(&H->G[pp], &p, phi + H->coff[pp], qb, fb);
         fscatter(H, pp, fb, flux, 1);
         free(fb);
         free(qb);
      }
      o_hier_coarsen_faces(H, l, flux, 1, H->nf);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* qb = q ? pgather(H, pp, q, ql) : NULL;
         double* rb = pgather(H, pp, rhs, 1);
         o_phase_rhs_fx(&H->G[pp], &p, flux + H->foff[pp], phi + H->coff[pp], qb,
                        cl ? cl + H->coff[pp] : NULL, cs ? cs + H->coff[pp] : NULL, rb);
         pscatter(H, pp, rb, rhs, 1);
         free(rb);
         free(qb);
      }
   }
   free(flux);
}

/* Face |grad q|^2 of every patch (side field, ghosts of q filled). */
void o_hier_quat_face_norm2(const OHier* H, const ampe_model_params* p0, const double* q,
                            double* n2)
{
   for (int l = 0; l < H->nlev; l++) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* qb = pgather(H, pp, q, p0->qlen);
         o_quat_face_norm2(&H->G[pp], &p, qb, n2 + H->foff[pp]);
         free(qb);
      }
   }
}

/* Composite orientation RHS with projection from lagged face norms n2:
 * finest to coarsest, the component fluxes fc dq/dx of the level, the finer
 * fluxes coarsened onto it, then the divergence and the projection
 * (QuatFACOps::evaluateRHS -> accumulateOperatorOnLevel with the outerside
 * flux coarsening, source/QuatFACOps.cc:1861-1890, 2040-2100). */
void o_hier_quat_rhs(const OHier* H, const ampe_model_params* p0, const double* phi,
                     const double* q, const double* n2, double* rhs)
{
   int ql = p0->qlen;
   double* flux = (double*)calloc((size_t)H->nf * ql, sizeof(double));
   for (int l = H->nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* qb = pgather(H, pp, q, ql);
         double* fb = (double*)malloc(sizeof(double) * pnf(&H->G[pp]) * ql);
         o_quat_flux(&H->G[pp], &p, phi + H->coff[pp], qb, n2 + H->foff[pp], fb);
         fscatter(H, pp, fb, flux, ql);
         free(fb);
         free(qb);
      }
      o_hier_coarsen_faces(H, l, flux, ql, H->nf);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* qb = pgather(H, pp, q, ql);
         double* fb = fgather(H, pp, flux, ql);
         double* rb = pgather(H, pp, rhs, ql);
         o_quat_rhs_fx(&H->G[pp], &p, phi + H->coff[pp], qb, fb, rb);
         pscatter(H, pp, rb, rhs, ql);
         free(rb);
         free(fb);
         free(qb);
      }
   }
   free(flux);
}

/* KKS composition coefficients D_c, D_phi on every patch's faces, then the
 * finer levels' coefficients coarsened onto the coarser faces
 * (QuatIntegrator::setDiffusionCoeffForConcentration, source/
 * QuatIntegrator.cc:2395-2428). */
void o_hier_conc_coeffs(const OHier* H, const ampe_model_params* p0, const double* phi,
                        const double* cl, const double* cs, double* Dc, double* Dphi)
{
   for (int l = 0; l < H->nlev; l++) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++)
         o_conc_coeffs(&H->G[pp], &p, phi + H->coff[pp], cl + H->coff[pp], cs + H->coff[pp],
                       Dc + H->foff[pp], Dphi + H->foff[pp]);
   }
   for (int l = H->nlev - 2; l >= 0; l--) {
      o_hier_coarsen_faces(H, l, Dc, 1, H->nf);
      o_hier_coarsen_faces(H, l, Dphi, 1, H->nf);
   }
}

/* Composite composition RHS, KKS form (QuatIntegrator::
 * evaluateConcentrationRHS, source/QuatIntegrator.cc:2578-2690). */
void o_hier_conc_rhs(const OHier* H, const ampe_model_params* p0, const double* c,
                     const double* phi, const double* Dc, const double* Dphi, double* rhs)
{
   double* flux = (double*)calloc((size_t)H->nf, sizeof(double));
   for (int l = H->nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++)
         o_conc_flux(&H->G[pp], &p, c + H->coff[pp], phi + H->coff[pp], Dc + H->foff[pp],
                     Dphi + H->foff[pp], flux + H->foff[pp]);
      o_hier_coarsen_faces(H, l, flux, 1, H->nf);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++)
         o_conc_rhs_fx(&H->G[pp], &p, flux + H->foff[pp], rhs + H->coff[pp]);
   }
   free(flux);
}

/* Composite three-phase multi-order phase RHS (ghosts of the orders filled on
 * every level, ph6 the phase concentrations of every cell): finest to
 * coarsest, every order's side fluxes of the level, the finer level's fluxes
 * coarsened onto it (the norder-deep flux coarsening of
 * MultiOrderRHSStrategy::evaluateRHS, source/MultiOrderRHSStrategy.cc:85-138),
 * then the RHS per patch. */
void o_hier_phase_rhs_tern3(const OHier* H, const ampe_model_params* p0, int norder,
                            const double* phi, const double* ph6, double* rhs)
{
   double* flux = (double*)calloc((size_t)H->nf * norder, sizeof(double));
   for (int l = H->nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* fb_phi = pgather(H, pp, phi, norder);
         double* fb = (double*)malloc(sizeof(double) * pnf(&H->G[pp]) * norder);
         o_phase_flux_multiorder(&H->G[pp], &p, norder, fb_phi, fb);
         fscatter(H, pp, fb, flux, norder);
         free(fb);
         free(fb_phi);
      }
      o_hier_coarsen_faces(H, l, flux, norder, H->nf);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* pb = pgather(H, pp, phi, norder);
         double* cb = pgather(H, pp, ph6, 6);
         double* fb = fgather(H, pp, flux, norder);
         double* rb = pgather(H, pp, rhs, norder);
         o_phase_rhs_tern3_fx(&H->G[pp], &p, norder, fb, pb, cb, NULL, rb);
         pscatter(H, pp, rb, rhs, norder);
         free(rb);
         free(fb);
         free(cb);
         free(pb);
      }
   }
   free(flux);
}

/* Composite multi-order phase RHS without composition (grain growth):
 * the flux coarsening of o_hier_phase_rhs_tern3, then computerhsmultiorder
 * per patch (MultiOrderRHSStrategy::evaluateRHS, source/
 * MultiOrderRHSStrategy.cc:85-138). */
void o_hier_phase_rhs_multiorder(const OHier* H, const ampe_model_params* p0, int norder,
                                 const double* phi, double* rhs)
{
   double* flux = (double*)calloc((size_t)H->nf * norder, sizeof(double));
   for (int l = H->nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* pb = pgather(H, pp, phi, norder);
         double* fb = (double*)malloc(sizeof(double) * pnf(&H->G[pp]) * norder);
         o_phase_flux_multiorder(&H->G[pp], &p, norder, pb, fb);
         fscatter(H, pp, fb, flux, norder);
         free(fb);
         free(pb);
      }
      o_hier_coarsen_faces(H, l, flux, norder, H->nf);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* pb = pgather(H, pp, phi, norder);
         double* fb = fgather(H, pp, flux, norder);
         double* rb = pgather(H, pp, rhs, norder);
         o_phase_rhs_multiorder_fx(&H->G[pp], &p, norder, fb, pb, NULL, rb);
         pscatter(H, pp, rb, rhs, norder);
         free(rb);
         free(fb);
         free(pb);
      }
   }
   free(flux);
}

/* Three-phase face diffusivities on every patch, then the finer levels'
 * coarsened onto the coarser faces (as o_hier_conc_coeffs). */
void o_hier_ebs3_coeffs(const OHier* H, const ampe_model_params* p0, int norder,
                        const double* phi, double* DL, double* DA, double* DB)
{
   for (int l = 0; l < H->nlev; l++) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* pb = pgather(H, pp, phi, norder);
         o_ebs3_coeffs(&H->G[pp], &p, norder, pb, DL + H->foff[pp], DA + H->foff[pp],
                       DB + H->foff[pp]);
         free(pb);
      }
   }
   for (int l = H->nlev - 2; l >= 0; l--) {
      o_hier_coarsen_faces(H, l, DL, 1, H->nf);
      o_hier_coarsen_faces(H, l, DA, 1, H->nf);
      o_hier_coarsen_faces(H, l, DB, 1, H->nf);
   }
}

/* Composite three-phase EBS composition RHS of the two species: finest to
 * coarsest, the level's fluxes, the finer fluxes coarsened onto it, then
 * M_c div J per patch (QuatIntegrator::evaluateConcentrationRHS,
 * source/QuatIntegrator.cc:2578-2690). */
void o_hier_conc_rhs_ebs3(const OHier* H, const ampe_model_params* p0, const double* ph6,
                          const double* DL, const double* DA, const double* DB, double* rhs)
{
   double* flux = (double*)calloc((size_t)H->nf * 2, sizeof(double));
   for (int l = H->nlev - 1; l >= 0; l--) {
      ampe_model_params p = lev_params(H, p0, l);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* cb = pgather(H, pp, ph6, 6);
         double* fb = (double*)malloc(sizeof(double) * pnf(&H->G[pp]) * 2);
         o_ebs3_flux(&H->G[pp], &p, cb, DL + H->foff[pp], DA + H->foff[pp], DB + H->foff[pp],
                     fb);
         fscatter(H, pp, fb, flux, 2);
         free(fb);
         free(cb);
      }
      o_hier_coarsen_faces(H, l, flux, 2, H->nf);
      for (int pp = H->pbeg[l]; pp < H->pbeg[l + 1]; pp++) {
         double* fb = fgather(H, pp, flux, 2);
         double* rb = pgather(H, pp, rhs, 2);
         o_conc_rhs_ebs3_fx(&H->G[pp], &p, fb, rb);
         pscatter(H, pp, rb, rhs, 2);
         free(rb);
         free(fb);
      }
   }
   free(flux);
}

/* ---------------------------------------------------------------- FAC */

/* the per-level operator of one preconditioner block */
static double blk_D(const OHBlock* B, long f) { return B->D ? B->D[f] : B->D_const; }

/* Side fluxes of level l for component m: D (w_c - w_m) / h with w = u
 * (elliptic blocks, efo_compfluxvardc / compfluxcondc) or w = S u (the
 * orientation block's symmetrised operator); then, on levels above the
 * coarsest, Ewing's fix of the coarse-fine faces (EllipticFACOps::
 * ewingFixFlux, source/EllipticFACOps.cc:1239; efo_ewingfixflux*).  The
 * reference's weights are single-precision constants: tranwt =
 * 1.0/(1+ratio) in REAL arithmetic, longwt = 2*tranwt. */
static void level_flux(const OHier* H, const OHBlock* B, int l, int lmin, const double* u,
                       double* flux)
{
   const double tranwt = (double)(1.0f / (float)(1 + 2)), longwt = 2.0 * tranwt;
   for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
      const OGeom* G = &H->G[p];
      long st[3] = {1, G->sy, G->sz};
      double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
      for (int d = 0; d < G->ndim; d++) {
         int ex = G->nx + (d == 0), ey = G->ny + (d == 1), ez = G->nz + (d == 2);
         for (int k = 0; k < ez; k++)
            for (int j = 0; j < ey; j++)
               for (int i = 0; i < ex; i++) {
                  long c0 = cidx(H, p, i, j, k), cm = c0 - st[d];
                  long fi = fidx(H, p, d, i, j, k);
                  double dd = blk_D(B, fi);
                  if (B->kind == 1)
                     flux[fi] = dinv[d] * dd * (B->S[c0] * u[c0] - B->S[cm] * u[cm]);
                  else
                     flux[fi] = dinv[d] * dd * (u[c0] - u[cm]);
               }
      }
      if (!B->ewing || l <= lmin) continue;
      /* coarse-fine faces: the outer cell is a coarse-fine ghost of this patch */
      for (int d = 0; d < G->ndim; d++)
         for (int side = 0; side < 2; side++) {
            int n[3] = {G->nx, G->ny, G->nz};
            int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            for (int b2 = 0; b2 < (G->ndim == 3 ? n[t2] : 1); b2++)
               for (int b1 = 0; b1 < n[t1]; b1++) {
                  int I[3] = {0, 0, 0}, gh[3];
                  I[d] = side ? n[d] - 1 : 0;
                  I[t1] = b1;
                  if (G->ndim == 3) I[t2] = b2;
                  for (int a = 0; a < 3; a++) gh[a] = I[a];
                  gh[d] += side ? 1 : -1;
                  int g[3] = {H->lo[p][0] + gh[0], H->lo[p][1] + gh[1], H->lo[p][2] + gh[2]};
                  /* outside the domain (not periodic): a physical boundary */
                  int outside = 0;
                  for (int a = 0; a < G->ndim; a++)
                     if (g[a] < 0 || g[a] >= H->dom[l][a]) outside = 1;
                  if (outside && !H->periodic) continue;
                  canon(H, l, g);
                  if (find(H, l, g) >= 0) continue;
                  /* the diametrically opposite fine cell of the coarse
                   * cell's face (jq = ratio - jp - 1) */
                  int Iq[3] = {I[0], I[1], I[2]};
                  Iq[t1] = I[t1] ^ 1;
                  if (G->ndim == 3) Iq[t2] = I[t2] ^ 1;
                  int F[3] = {I[0], I[1], I[2]};
                  F[d] = side ? n[d] : 0;
                  long fi = fidx(H, p, d, F[0], F[1], F[2]);
                  long ci = cidx(H, p, I[0], I[1], I[2]);
                  long cq = cidx(H, p, Iq[0], Iq[1], Iq[2]);
                  double dd = blk_D(B, fi);
                  double h = G->dx[d];
                  if (side == 0)
                     flux[fi] = longwt * flux[fi] + tranwt * dd * (u[cq] - u[ci]) / h;
                  else
                     flux[fi] = longwt * flux[fi] - tranwt * dd * (u[cq] - u[ci]) / h;
               }
         }
   }
}

static double cell_div(const OHier* H, int p, const double* flux, int i, int j, int k)
{
   const OGeom* G = &H->G[p];
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   double s = dinv[0] * (flux[fidx(H, p, 0, i + 1, j, k)] - flux[fidx(H, p, 0, i, j, k)]) +
              dinv[1] * (flux[fidx(H, p, 1, i, j + 1, k)] - flux[fidx(H, p, 1, i, j, k)]);
   if (G->ndim == 3)
      s = s + dinv[2] * (flux[fidx(H, p, 2, i, j, k + 1)] - flux[fidx(H, p, 2, i, j, k)]);
   return s;
}

static double cell_dsum(const OHier* H, const OHBlock* B, int p, int i, int j, int k)
{
   const OGeom* G = &H->G[p];
   double dinv[3] = {1.0 / G->dx[0], 1.0 / G->dx[1], 1.0 / G->dx[2]};
   double s = dinv[0] * dinv[0] *
                  (blk_D(B, fidx(H, p, 0, i + 1, j, k)) + blk_D(B, fidx(H, p, 0, i, j, k))) +
              dinv[1] * dinv[1] *
                  (blk_D(B, fidx(H, p, 1, i, j + 1, k)) + blk_D(B, fidx(H, p, 1, i, j, k)));
   if (G->ndim == 3)
      s = s + dinv[2] * dinv[2] *
                  (blk_D(B, fidx(H, p, 2, i, j, k + 1)) + blk_D(B, fidx(H, p, 2, i, j, k)));
   return s;
}

/* res = f - A u at one cell from the flux: elliptic M div F + C u
 * (efo_compres*sca, source/fortran/2d/ellipticfacops.m4:346-446); the
 * orientation block u + gamma S div F */
static double cell_res(const OHier* H, const OHBlock* B, int p, const double* f,
                       const double* u, const double* flux, int i, int j, int k)
{
   long c0 = cidx(H, p, i, j, k);
   double div = cell_div(H, p, flux, i, j, k);
   if (B->kind == 1) return f[c0] - B->gamma * B->S[c0] * div - 1.0 * u[c0];
   double m = B->M ? B->M[c0] : B->M_const;
   double cc = B->C ? B->C[c0] : B->C_const;
   return f[c0] - m * div - cc * u[c0];
}

/* one colour of red-black Gauss-Seidel on level l from the flux (computed
 * before the colour); returns max |res| */
static double level_rbgs_color(const OHier* H, const OHBlock* B, int l, const double* f,
                               double* u, const double* flux, int color)
{
   double maxres = 0.0;
   for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
      const OGeom* G = &H->G[p];
      for (int k = 0; k < G->nz; k++)
         for (int j = 0; j < G->ny; j++) {
            int gj = H->lo[p][1] + j;
            int x = (G->ndim == 3) ? H->lo[p][0] + gj + (H->lo[p][2] + k) : H->lo[p][1] + gj;
            int io = ((x - (x / 2) * 2) != color) ? 1 : 0;
            for (int i = io; i < G->nx; i += 2) {
               long c0 = cidx(H, p, i, j, k);
               double res = cell_res(H, B, p, f, u, flux, i, j, k);
               double dsum = cell_dsum(H, B, p, i, j, k);
               if (B->kind == 1) {
                  double center = 1.0 - B->gamma * B->S[c0] * (B->S[c0] * dsum);
                  u[c0] = u[c0] + res / center;
               } else {
                  double m = B->M ? B->M[c0] : B->M_const;
                  double cc = B->C ? B->C[c0] : B->C_const;
                  double dudr = 1.0 / (m * dsum - cc);
                  double du = -res * dudr;
                  u[c0] = u[c0] + du * 1.0;
               }
               if (maxres < fabs(res)) maxres = fabs(res);
            }
         }
   }
   return maxres;
}

typedef struct FacWork {
   double* flux;    /* level fluxes, one component */
   double* oflux;   /* saved fluxes of the finer levels */
   double* tmp;     /* residual of the cycle (tmp residual vector) */
} FacWork;

/* EllipticFACOps::smoothErrorByRedBlack (source/EllipticFACOps.cc:1067-1230):
 * ghosts of the level without the coarse level, a one-time coarse-fine fill
 * (CONSTANT_REFINE) above the coarsest level, sweeps of red then black with
 * a same-level fill before each colour, early exit when tol >= 0 */
static void level_smooth(const OHier* H, const OHBlock* B, int l, int lmin, const double* f,
                         double* u, int ncomp, int sweeps, double tol, FacWork* W)
{
   for (int m = 0; m < ncomp; m++) {
      double* um = u + (long)m * H->ncg;
      o_hier_ghost_fill_level(H, l, um, 1, 0);
      if (l > lmin) o_hier_ghost_fill_level(H, l, um, 1, 1);
   }
   int not_converged = 1;
   for (int s = 0; s < sweeps && not_converged; s++) {
      double rmax = 0.0, bmax = 0.0;
      for (int m = 0; m < ncomp; m++) {
         double* um = u + (long)m * H->ncg;
         const double* fm = f + (long)m * H->ncg;
         o_hier_ghost_fill_level(H, l, um, 1, 0);
         level_flux(H, B, l, lmin, um, W->flux);
         double r = level_rbgs_color(H, B, l, fm, um, W->flux, 0);
         o_hier_ghost_fill_level(H, l, um, 1, 0);
         level_flux(H, B, l, lmin, um, W->flux);
         double b = level_rbgs_color(H, B, l, fm, um, W->flux, 1);
         o_hier_ghost_fill_level(H, l, um, 1, 0);
         if (r > rmax) rmax = r;
         if (b > bmax) bmax = b;
      }
      if (tol >= 0.0) not_converged = (rmax > bmax ? rmax : bmax) > tol;
   }
}

/* EllipticFACOps::computeCompositeResidualOnLevel (:1480-1600): ghosts
 * (coarse-fine by CONSTANT_REFINE above the coarsest level), fluxes with
 * Ewing's fix, the finer level's saved outer fluxes coarsened onto this
 * level, the residual, and this level's fluxes saved for the next coarser */
static void level_comp_residual(const OHier* H, const OHBlock* B, int l, int lmin, int lmax,
                                double* res, double* u, const double* f, int ncomp,
                                FacWork* W)
{
   for (int m = 0; m < ncomp; m++) {
      double* um = u + (long)m * H->ncg;
      const double* fm = f + (long)m * H->ncg;
      double* rm = res + (long)m * H->ncg;
      double* om = W->oflux + (long)m * H->nf;
      o_hier_ghost_fill_level(H, l, um, 1, l > lmin ? 1 : 0);
      level_flux(H, B, l, lmin, um, W->flux);
      if (l < lmax) coarsen_faces(H, l, om, W->flux, 1, H->nf, 1);
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         const OGeom* G = &H->G[p];
         for (int k = 0; k < G->nz; k++)
            for (int j = 0; j < G->ny; j++)
               for (int i = 0; i < G->nx; i++)
                  rm[cidx(H, p, i, j, k)] = cell_res(H, B, p, fm, um, W->flux, i, j, k);
         if (l > lmin)
            memcpy(om + H->foff[p], W->flux + H->foff[p], sizeof(double) * pnf(G));
      }
   }
}

/* prolongErrorAndCorrect with CONSTANT_REFINE (:990-1041): e_l += parent value */
static void level_prolong_add(const OHier* H, int l, double* e, int ncomp)
{
   for (int m = 0; m < ncomp; m++) {
      double* a = e + (long)m * H->ncg;
      for (int p = H->pbeg[l]; p < H->pbeg[l + 1]; p++) {
         const OGeom* G = &H->G[p];
         for (int k = 0; k < G->nz; k++)
            for (int j = 0; j < G->ny; j++)
               for (int i = 0; i < G->nx; i++) {
                  int g[3] = {fdiv2(H->lo[p][0] + i), fdiv2(H->lo[p][1] + j),
                              H->ndim == 3 ? fdiv2(H->lo[p][2] + k) : 0};
                  long c = find(H, l - 1, g);
                  long ci = cidx(H, p, i, j, k);
                  a[ci] = a[ci] + a[c];
               }
      }
   }
}

static double rms_norm(const OHier* H, const double* r, int ncomp, const double* w)
{
   double s = 0.0, sw = 0.0;
   for (int m = 0; m < ncomp; m++)
      for (int p = 0; p < H->np; p++) {
         const OGeom* G = &H->G[p];
         for (int k = 0; k < G->nz; k++)
            for (int j = 0; j < G->ny; j++)
               for (int i = 0; i < G->nx; i++) {
                  long c = cidx(H, p, i, j, k);
                  double v = r[(long)m * H->ncg + c];
                  s += w[c] * v * v;
                  sw += w[c];
               }
      }
   return sqrt(s / sw);
}

/* FACPreconditioner::computeFullCompositeResidual (:661-680) */
static double full_comp_residual(const OHier* H, const OHBlock* B, double* r, double* u,
                                 const double* f, int ncomp, const double* w, FacWork* W)
{
   int lmax = H->nlev - 1;
   level_comp_residual(H, B, lmax, 0, lmax, r, u, f, ncomp, W);
   for (int l = lmax - 1; l >= 0; l--) {
      o_hier_coarsen_level(H, l, u, ncomp, 0);
      level_comp_residual(H, B, l, 0, lmax, r, u, f, ncomp, W);
      o_hier_coarsen_level(H, l, r, ncomp, 0);
   }
   return rms_norm(H, r, ncomp, w);
}

/* FACPreconditioner::facCycle_Recursive (:431-474) */
static void fac_cycle(const OHier* H, const OHBlock* B, const OHFacParams* P, double* e,
                      double* r, int ncomp, int l, FacWork* W)
{
   int lmax = H->nlev - 1;
   if (l == 0) {
      /* solveCoarsestLevel, "redblack" (:1356-1385) */
      level_smooth(H, B, 0, 0, r, e, ncomp, P->coarse_its, B->kind == 0 ? P->coarse_tol : -1.0,
                   W);
      for (int m = 0; m < ncomp; m++) o_hier_ghost_fill_level(H, 0, e + (long)m * H->ncg, 1, 0);
      return;
   }
   level_smooth(H, B, l, 0, r, e, ncomp, P->pre, -1.0, W);
   level_comp_residual(H, B, l, 0, lmax, W->tmp, e, r, ncomp, W);
   /* restrictResidual: the covered coarse cells of r take the fine residual */
   {
      for (int m = 0; m < ncomp; m++) {
         /* coarsen tmp (level l) into r (level l-1) */
         double* tm = W->tmp + (long)m * H->ncg;
         double* rm = r + (long)m * H->ncg;
         double* save = (double*)malloc(sizeof(double) * H->ncg);
         memcpy(save, tm, sizeof(double) * H->ncg);
         /* copy the level l-1 entries of r into tmp, coarsen, copy back */
         for (int p = H->pbeg[l - 1]; p < H->pbeg[l]; p++)
            memcpy(tm + H->coff[p], rm + H->coff[p], sizeof(double) * H->G[p].ncg);
         o_hier_coarsen_level(H, l - 1, tm, 1, 0);
         for (int p = H->pbeg[l - 1]; p < H->pbeg[l]; p++)
            memcpy(rm + H->coff[p], tm + H->coff[p], sizeof(double) * H->G[p].ncg);
         memcpy(tm, save, sizeof(double) * H->ncg);
         free(save);
      }
   }
   fac_cycle(H, B, P, e, r, ncomp, l - 1, W);
   level_prolong_add(H, l, e, ncomp);
   level_smooth(H, B, l, 0, r, e, ncomp, P->post, -1.0, W);
}

/* the full composite residual r = f - A u over all levels and its weighted
 * RMS norm (FACPreconditioner::computeFullCompositeResidual) */
double o_hier_comp_residual(const OHier* H, const OHBlock* B, double* u, const double* f,
                            double* r, int ncomp)
{
   double* w = (double*)malloc(sizeof(double) * H->ncg);
   FacWork W;
   W.flux = (double*)calloc((size_t)H->nf, sizeof(double));
   W.oflux = (double*)calloc((size_t)H->nf * ncomp, sizeof(double));
   W.tmp = NULL;
   o_hier_weights(H, w);
   double nrm = full_comp_residual(H, B, r, u, f, ncomp, w, &W);
   free(w);
   free(W.flux);
   free(W.oflux);
   return nrm;
}

/* FACPreconditioner::solveSystem (:267-429): u = 0 on entry is the caller's
 * choice; returns 0 when the RMS composite residual ends below tol (1 not),
 * *norm and *cycles (may be NULL) report the last norm and the cycles. */
int o_hier_fac_solve(const OHier* H, const OHBlock* B, const OHFacParams* P, const double* f,
                     double* u, int ncomp, double tol, double* norm, int* cycles)
{
   long n = H->ncg * ncomp;
   double* e = (double*)calloc((size_t)n, sizeof(double));
   double* r = (double*)calloc((size_t)n, sizeof(double));
   double* w = (double*)malloc(sizeof(double) * H->ncg);
   FacWork W;
   W.flux = (double*)calloc((size_t)H->nf, sizeof(double));
   W.oflux = (double*)calloc((size_t)H->nf * ncomp, sizeof(double));
   W.tmp = (double*)calloc((size_t)n, sizeof(double));
   o_hier_weights(H, w);
   double nrm = full_comp_residual(H, B, r, u, f, ncomp, w, &W);
   int it = 0;
   do {
      memset(e, 0, sizeof(double) * n);
      fac_cycle(H, B, P, e, r, ncomp, H->nlev - 1, &W);
      for (int m = 0; m < ncomp; m++)
         for (int p = 0; p < H->np; p++) {
            const OGeom* G = &H->G[p];
            for (int k = 0; k < G->nz; k++)
               for (int j = 0; j < G->ny; j++)
                  for (int i = 0; i < G->nx; i++) {
                     long c = (long)m * H->ncg + cidx(H, p, i, j, k);
                     u[c] = u[c] + e[c];
                  }
         }
      for (int l = H->nlev - 2; l >= 0; l--) o_hier_coarsen_level(H, l, u, ncomp, 0);
      nrm = full_comp_residual(H, B, r, u, f, ncomp, w, &W);
      it++;
   } while (nrm > tol && it < P->max_cycles);
   if (norm) *norm = nrm;
   if (cycles) *cycles = it;
   free(e);
   free(r);
   free(w);
   free(W.flux);
   free(W.oflux);
   free(W.tmp);
   return nrm < tol ? 0 : 1;
}
