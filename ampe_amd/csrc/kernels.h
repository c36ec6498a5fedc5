// Device-side indexing helpers and the launch interface of the RHS, FAC and
// vector kernels.
#pragma once

#include "level.h"
#include "physics.h"

namespace ampe {

// Everything a level kernel needs to locate cells and faces.
struct LevelArgs {
   const PatchDev* P;
   const Tile* T;
   int g, gz;
   long long ncg;
   long long sdo[3];  // direction-block offsets inside a side field
   int nz_tile;       // planes per tile (1, or KZ for column tiles)
   const double* wt;  // cell weights of the reductions (null: 1)
};


inline LevelArgs level_args(const Level& L, bool faces = false)
{
   LevelArgs a;
   a.P = L.dp;
   a.T = faces ? L.face_tiles : L.cell_tiles;
   a.nz_tile = 1;
   a.g = L.g;
   a.gz = L.gz;
   a.ncg = L.ncg;
   for (int d = 0; d < 3; d++) a.sdo[d] = L.side_dir_off[d];
   a.wt = L.wt;
   return a;
}

// Column tile list (KZ planes per workgroup in 3-D).
inline LevelArgs level_args_col(const Level& L)
{
   LevelArgs a = level_args(L);
   a.T = L.col_tiles;
   a.nz_tile = L.ndim == 3 ? KZ : 1;
   return a;
}

// Column list of the marching kernels (KM planes per workgroup in 3-D).
inline LevelArgs level_args_march(const Level& L)
{
   LevelArgs a = level_args(L);
   a.T = L.march_tiles;
   a.nz_tile = L.ndim == 3 ? KM : 1;
   return a;
}

struct CellPos {
   int i, j, k;
   long long c0;  // ghosted cell index inside one component block
   long long st[3];
   PatchDev pd;
};

// Tile of this workgroup: the tile lists are stored in dispatch order, already
// arranged per XCD on the host (level_create).
__device__ __forceinline__ int tile_index(const LevelArgs&) { return (int)blockIdx.x; }

// Cell of this thread in the tile list; false when outside the box
// (extend = 1 for face tiles: boxes grown by one in each direction).
__device__ __forceinline__ bool cell_pos(const LevelArgs& a, int extend, CellPos& c)
{
   const Tile t = a.T[tile_index(a)];
   c.pd = a.P[t.patch];
   c.i = t.i0 + (int)threadIdx.x;
   c.j = t.j0 + (int)threadIdx.y;
   c.k = t.k;
   if (c.i >= c.pd.nx + extend || c.j >= c.pd.ny + extend) return false;
   c.st[0] = 1;
   c.st[1] = c.pd.sy;
   c.st[2] = c.pd.sz;
   c.c0 = c.pd.cell_off + (c.i + a.g) + (long long)(c.j + a.g) * c.pd.sy +
          (long long)(c.k + a.gz) * c.pd.sz;
   return true;
}

// Cell of this thread at plane t.k + kk of a column tile; false when outside
// the box (i, j or k).
__device__ __forceinline__ bool cell_pos_col(const LevelArgs& a, int kk, CellPos& c)
{
   const Tile t = a.T[tile_index(a)];
   c.pd = a.P[t.patch];
   c.i = t.i0 + (int)threadIdx.x;
   c.j = t.j0 + (int)threadIdx.y;
   c.k = t.k + kk;
   if (c.i >= c.pd.nx || c.j >= c.pd.ny || c.k >= c.pd.nz) return false;
   c.st[0] = 1;
   c.st[1] = c.pd.sy;
   c.st[2] = c.pd.sz;
   c.c0 = c.pd.cell_off + (c.i + a.g) + (long long)(c.j + a.g) * c.pd.sy +
          (long long)(c.k + a.gz) * c.pd.sz;
   return true;
}

// Pair mapping of a tile onto a workgroup of TX/2 x TY threads: thread
// (tx, ty) takes the cell 2 tx + io of its row, so the lanes of a half-wave
// hold cells of one column parity (io = 0 and 1 in turn cover the tile).
// Plane t.k + kk; false when outside the box.
__device__ __forceinline__ bool cell_pos_pair(const LevelArgs& a, int kk, int io, CellPos& c)
{
   const Tile t = a.T[tile_index(a)];
   c.pd = a.P[t.patch];
   c.i = t.i0 + 2 * (int)threadIdx.x + io;  // t.i0 is a multiple of TX (even)
   c.j = t.j0 + (int)threadIdx.y;
   c.k = t.k + kk;
   if (c.i >= c.pd.nx || c.j >= c.pd.ny || c.k >= c.pd.nz) return false;
   c.st[0] = 1;
   c.st[1] = c.pd.sy;
   c.st[2] = c.pd.sz;
   c.c0 = c.pd.cell_off + (c.i + a.g) + (long long)(c.j + a.g) * c.pd.sy +
          (long long)(c.k + a.gz) * c.pd.sz;
   return true;
}

// Row parity-split layout of a ghosted cell block: inside every row of sy
// cells the even ghosted columns come first, then the odd ones, so the cells
// of one red-black colour in a row are contiguous (a colour sweep writes and
// reads its own cells in whole lines).  Rows, planes and patches stay where
// the natural layout has them; only the position inside a row changes.
// Position of ghosted column ig in its row:
__host__ __device__ __forceinline__ int split_col(int ig, int sy)
{
   return (ig & 1) ? ((sy + 1) >> 1) + (ig >> 1) : (ig >> 1);
}

// Split-layout index of the cell di columns (-1, 0, 1) beside c.
__device__ __forceinline__ long long split_at(const LevelArgs& a, const CellPos& c, int di)
{
   const int ig = c.i + a.g;
   return c.c0 - ig + split_col(ig + di, c.pd.sy);
}

// Split-layout index of ghosted cell n of one component block (natural
// index), the patch found by bisection over the ascending cell offsets.
__device__ __forceinline__ long long split_index(const PatchDev* __restrict__ P, int np,
                                                 long long n)
{
   int lo = 0, hi = np - 1;
   while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (P[mid].cell_off <= n)
         lo = mid;
      else
         hi = mid - 1;
   }
   const long long loc = n - P[lo].cell_off;
   const int sy = P[lo].sy;
   const int ig = (int)(loc % sy);
   return n - ig + split_col(ig, sy);
}

// Per-plane loop of a column kernel: the body runs once per plane of the
// thread's column, with `c` the cell; `continue` inside the body moves to
// the next plane.
#define AMPE_COLUMN_BEGIN(a, c)                      \
   for (int kk_ = 0; kk_ < (a).nz_tile; kk_++) {     \
      CellPos c;                                     \
      if (!cell_pos_col((a), kk_, c)) break;         \
      do {
#define AMPE_COLUMN_END \
   }                    \
   while (0)            \
      ;                 \
   }

// Launch-uniform constants of the right-hand-side kernels (built on the host
// from the model parameters by rhs_const).
struct RhsConst {
   double dxinv[3];   // 1/dx
   double hdxinv[3];  // 0.5/dx
   double e2dx[3];    // eps_phi^2/dx
   double e2dx12[3];  // eps_phi^2/(12 dx)
   double convL, convA;  // 1e-6/v_m of the liquid and the solid
   double maxgni;     // 1/gradient_floor
   double Dl, Ds;     // D_L0 exp(-Q_L/RT), D_S0 exp(-Q_S/RT)
};
RhsConst rhs_const(const ampe_model_params& p);

// Index of face (i,j,k) of direction d of patch pd in a side field.
__device__ __forceinline__ long long face_idx(const LevelArgs& a, const PatchDev& pd,
                                              int d, int i, int j, int k)
{
   int ex = pd.nx + (d == 0), ey = pd.ny + (d == 1);
   return a.sdo[d] + pd.side_off[d] + i + (long long)ex * (j + (long long)ey * k);
}

// Weight of cell c0 in a reduction (1 without weights).
__device__ __forceinline__ double red_w(const LevelArgs& a, long long c0)
{
   return a.wt ? a.wt[c0] : 1.0;
}

// Compact interior index (host I/O layout).
__device__ __forceinline__ long long int_idx(const CellPos& c)
{
   return c.pd.int_off + c.i + (long long)c.pd.nx * (c.j + (long long)c.pd.ny * c.k);
}

// ------------------------------------------------------------ launchers
int launch_quat_grad(const Level& L, const ampe_model_params& p, const double* q,
                     double* gmod, double* face_norm2, hipStream_t s);
int launch_phase_rhs(const Level& L, const ampe_model_params& p, const double* phi,
                     const double* q, const double* cl, const double* cs, double* rhs,
                     hipStream_t s);
int launch_quat_rhs(const Level& L, const ampe_model_params& p, const double* phi,
                    const double* q, const double* n2, double* rhs, hipStream_t s);
// The same from face coefficients computed beforehand by launch_quat_face_coef.
int launch_quat_rhs_fc(const Level& L, const ampe_model_params& p, const double* phi,
                       const double* q, const double* fc, double* rhs, hipStream_t s);
int launch_quat_project(const Level& L, const ampe_model_params& p, const double* q,
                        double* corr, double* err, hipStream_t s);
int launch_conc_coeffs(const Level& L, const ampe_model_params& p, const double* phi,
                       const double* cl, const double* cs, double* Dc, double* Dphi,
                       hipStream_t s);
int launch_conc_rhs(const Level& L, const ampe_model_params& p, const double* c,
                    const double* phi, const double* Dc, const double* Dphi, double* rhs,
                    hipStream_t s);
// EBS face diffusivities of the composition-dependent (CALPHAD mobility) model.
int launch_ebs_coeffs_mob(const Level& L, const ampe_model_params& p, const double* phi,
                          const double* cl, const double* cs, double* DL, double* DA,
                          hipStream_t s);
int launch_ebs_coeffs(const Level& L, const ampe_model_params& p, const double* phi, double* DL,
                      double* DA, hipStream_t s);
int launch_conc_rhs_ebs(const Level& L, const ampe_model_params& p, const double* cl,
                        const double* cs, const double* DL, const double* DA, double* rhs,
                        hipStream_t s);
int launch_phase_rhs_multiorder(const Level& L, const ampe_model_params& p, int norder,
                                const double* phi, const double* mobility, double* rhs,
                                hipStream_t s);
int launch_project_phases(const Level& L, int norder, const double* phi, double* corr,
                          double* err, hipStream_t s);
// three phases on norderp > 1 orders, ternary quadratic (ternary.hip); ph:
// six ghosted fields cL0, cL1, cA0, cA1, cB0, cB1; c and the composition rhs:
// two species
int launch_kks_tern3(const Level& L, const ampe_model_params& p, int norder, const double* phi,
                     const double* c, double* ph, hipStream_t s);
int launch_phase_rhs_tern3(const Level& L, const ampe_model_params& p, int norder,
                           const double* phi, const double* ph, const double* mobility,
                           double* rhs, hipStream_t s);
int launch_ebs3_coeffs(const Level& L, const ampe_model_params& p, int norder, const double* phi,
                       double* DL, double* DA, double* DB, hipStream_t s);
int launch_conc_rhs_ebs3(const Level& L, const ampe_model_params& p, const double* ph,
                         const double* DL, const double* DA, const double* DB, double* rhs,
                         hipStream_t s);
int launch_conc_rhs_ebs_4th(const Level& L, const ampe_model_params& p, const double* cl,
                            const double* cs, const double* DL, const double* DA, double* rhs,
                            hipStream_t s);
int launch_conc_rhs_ebs_iso(const Level& L, const ampe_model_params& p, const double* phid,
                            const double* cl, const double* cs, double* rhs, hipStream_t s);
int launch_phase_mobility(const Level& L, const ampe_model_params& p, const double* cl,
                          const double* cs, double* M, hipStream_t s);
// Grain labels of the level's interior cells (interior-compact order, -1
// outside grains); synchronises s.
int grains_label(const Level& L, const double* phase, double thr, int* labels, int* ngrains,
                 int* passes, hipStream_t s);
int launch_tag(const Level& L, const ampe_model_params& p, const double* phi, double tol,
               int* tags, hipStream_t s);
// Berger-Rigoutsos clustering of a host mask over the box dom, or of a
// level's device tags (cluster.hip).
int cluster_mask(int ndim, const ampe_box& dom, const unsigned char* mask,
                 const ampe_cluster_params& cp, ampe_box* boxes, int max_boxes, int* nboxes);
int cluster_tags(const Level& L, const int* tags, const ampe_cluster_params& cp,
                 ampe_box* boxes, int max_boxes, int* nboxes, hipStream_t s);
int launch_quat_face_coef(const Level& L, const ampe_model_params& p, const double* phi,
                          const double* n2, double* fc, hipStream_t s);
int launch_sqrt_quat_mobility(const Level& L, const ampe_model_params& p, const double* phi,
                              double* sm, hipStream_t s);

// ------------------------------------------------------------ hierarchy
// A coarse-fine face of a fine patch: the face (composite side index), the
// cells below and above it along d (composite indices; one is the patch's
// ghost), the patch's strides, the ghost's coarse parent and the interior
// cell's mirror across its coarse cell (Ewing's diametrically opposite cell).
struct CfFace {
   long long face, cm, c0, sy, sz, parent, mirror;
   int d, side;  // side 0: the patch's lower side along d (the ghost is cm)
};
// A coarse cell next to coarse-fine faces: for each such face, the coarse
// face, its two cells and the first of its 2^(ndim-1) fine faces in the
// finer level's CfFace list, in the order of the side-data coarsening.
struct CfCorr {
   long long cell, sy, sz;
   long long face[6], cm[6], c0[6];
   int d[6], side[6], fine[6];
   int nface;
};
// Composite right-hand sides: the coarse level's fused kernel has run; each
// listed coarse cell's divergence takes the finer level's fluxes (their
// area-weighted average, af / ac) on its coarse-fine faces.
int launch_hier_phase_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                           const ampe_model_params& pc, const ampe_model_params& pf,
                           long long ncg, const double* phi, const double* q, const double* cl,
                           const double* cs, double af, double ac, double* rhs, hipStream_t s);
// the three-phase multi-order model: every order's phase flux and both
// species' EBS fluxes (phase mobility constant)
int launch_hier_tern3_phase_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                                 const ampe_model_params& pc, const ampe_model_params& pf,
                                 int norder, long long ncg, const double* phi, double af,
                                 double ac, double* rhs, hipStream_t s);
int launch_hier_ebs3_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                          const ampe_model_params& pc, const ampe_model_params& pf, long long ncg,
                          const double* ph, const double* DL, const double* DA, const double* DB,
                          double af, double ac, double* rhs, hipStream_t s);
int launch_hier_conc_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                          const ampe_model_params& pc, const ampe_model_params& pf,
                          const double* c, const double* phi, const double* Dc,
                          const double* Dphi, double af, double ac, double* rhs, hipStream_t s);
int launch_hier_quat_corr(int ndim, const CfCorr* cc, int ncorr, const CfFace* ff,
                          const ampe_model_params& pc, const ampe_model_params& pf,
                          long long ncg, const double* phi, const double* q, const double* fc,
                          double af, double ac, double* rhs, hipStream_t s);

}  // namespace ampe
