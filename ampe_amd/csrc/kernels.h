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
};

inline LevelArgs level_args(const Level& L, bool faces = false)
{
   LevelArgs a;
   a.P = L.dp;
   a.T = faces ? L.face_tiles : L.cell_tiles;
   a.g = L.g;
   a.gz = L.gz;
   a.ncg = L.ncg;
   for (int d = 0; d < 3; d++) a.sdo[d] = L.side_dir_off[d];
   return a;
}

struct CellPos {
   int i, j, k;
   long long c0;  // ghosted cell index inside one component block
   long long st[3];
   PatchDev pd;
};

// Cell of this thread in the tile list; false when outside the box
// (extend = 1 for face tiles: boxes grown by one in each direction).
__device__ __forceinline__ bool cell_pos(const LevelArgs& a, int extend, CellPos& c)
{
   const Tile t = a.T[blockIdx.x];
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

// Index of face (i,j,k) of direction d of patch pd in a side field.
__device__ __forceinline__ long long face_idx(const LevelArgs& a, const PatchDev& pd,
                                              int d, int i, int j, int k)
{
   int ex = pd.nx + (d == 0), ey = pd.ny + (d == 1);
   return a.sdo[d] + pd.side_off[d] + i + (long long)ex * (j + (long long)ey * k);
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
int launch_quat_project(const Level& L, const ampe_model_params& p, const double* q,
                        double* corr, double* err, hipStream_t s);
int launch_conc_coeffs(const Level& L, const ampe_model_params& p, const double* phi,
                       const double* cl, const double* cs, double* Dc, double* Dphi,
                       hipStream_t s);
int launch_conc_rhs(const Level& L, const ampe_model_params& p, const double* c,
                    const double* phi, const double* Dc, const double* Dphi, double* rhs,
                    hipStream_t s);
int launch_tag(const Level& L, const ampe_model_params& p, const double* phi, double tol,
               int* tags, hipStream_t s);
int launch_quat_face_coef(const Level& L, const ampe_model_params& p, const double* phi,
                          const double* n2, double* fc, hipStream_t s);
int launch_sqrt_quat_mobility(const Level& L, const ampe_model_params& p, const double* phi,
                              double* sm, hipStream_t s);

}  // namespace ampe
