// Geometric multigrid below one patch level, the preconditioner's alternative
// to the single-level red-black coarse solve (the role hypre PFMG plays as
// EllipticFACOps' coarse solver, source/EllipticFACOps.cc:1356-1385, and in
// QuatLevelSolver, source/QuatLevelSolver.cc:1210-1339).
//
// Coarse levels halve every patch box of the fine level (same patch list,
// same ghost width), so ghost filling, tiles and the red-black / residual
// kernels are those of the fine level.  Restriction: average of the 2^d
// children (cell fields) or of the 2^(d-1) fine faces covering a coarse face
// (face coefficients); prolongation: cell-centred tri/bilinear interpolation
// added to the finer correction.  The coarse operators are rediscretised with
// restricted coefficients and doubled mesh spacing.
#pragma once

#include <vector>

#include "fac.h"

namespace ampe {

template <typename T>
struct MgLevel {
   Level* L = nullptr;  // owned for levels >= 1
   double dx[3] = {1, 1, 1};
   T* C = nullptr;      // phase block C (cell), levels >= 1
   T* M = nullptr;      // phase block mobility (cell), levels >= 1
   T* D = nullptr;      // composition block D (side), levels >= 1
   T* fc = nullptr;     // orientation face coefficients (side), levels >= 1
   T* sm = nullptr;     // orientation sqrt(M) (cell, ghosted), levels >= 1
   T* x = nullptr;      // correction, ncomp_max blocks, levels >= 1
   T* b = nullptr;      // right-hand side, ncomp_max blocks, levels >= 1
   T* r = nullptr;      // residual, ncomp_max blocks, all levels
   T* t = nullptr;      // second buffer of the fused orientation sweeps, all levels
};

template <typename T>
struct Mg {
   std::vector<MgLevel<T>> lv;  // lv[0] is the fine level (not owned)
   int ncomp_max = 1;
   bool qsplit = false;         // orientation vectors in the row parity-split layout
   int nlev() const { return (int)lv.size(); }
};

struct MgCycle {
   int nu1 = 2, nu2 = 2;  // pre / post red-black sweeps
   int ncoarse = 16;      // sweeps on the coarsest level
   // orientation block: one-pass sweeps (same values; measured slower on the
   // 256^3 workload: 0.32 ms per sweep against 2 x 0.10 ms, averaged over the
   // levels -- each workgroup walks its column of planes with two barriers per
   // plane, which leaves too few loads in flight)
   bool fused = false;
};

// Build the hierarchy below fine while every patch box and the periodic
// domain have even lower corners and extents and the halved boxes keep at
// least min_cells cells per direction; at most max_levels levels in total.
template <typename T>
int mg_create(const Level& fine, const double* dx, int ncomp_max, int max_levels, int min_cells,
              Mg<T>** out);
template <typename T>
void mg_destroy(Mg<T>* m);

// Coarse-level coefficients restricted from fine-level arrays.
template <typename T>
int mg_setup_phase(Mg<T>& m, const T* C0, const T* M0, hipStream_t s);
template <typename T>
int mg_setup_conc(Mg<T>& m, const T* D0, hipStream_t s);
template <typename T>
int mg_setup_quat(Mg<T>& m, const T* fc0, const T* sm0, hipStream_t s);

// One V-cycle on A x = b from x = 0 (x0 zero on entry), fine arrays b0, x0.
// Elliptic blocks: kind 0 = phase (C field, D constant, M constant or field), kind 1 =
// composition (D field, C and M constants); op0 is the fine operator.
template <typename T>
int mg_vcycle_ell(Mg<T>& m, const EllOp<T>& op0, int kind, const MgCycle& cy, const T* b0,
                  T* x0, hipStream_t s);
// Orientation block, ncomp components sharing fc / sqrt(M).
template <typename T>
int mg_vcycle_quat(Mg<T>& m, const T* fc0, const T* sm0, double gamma, int ncomp,
                   const MgCycle& cy, const T* b0, T* x0, hipStream_t s);

// Store the orientation block's vectors (b, x, r of every level, and the
// fine-level b0 / x0 the caller passes) in the row parity-split layout
// (split_col): one colour of a sweep then reads and writes whole lines.
// Builds the levels' split ghost gathers; the one-pass sweeps are natural
// only.
template <typename T>
int mg_enable_qsplit(Mg<T>& m, hipStream_t s);

// Transfer operators (exposed for tests through the C ABI); split: f / c, xc
// / xf in the row parity-split layout.
template <typename T>
int mg_restrict_cell(const Level& coarse, const Level& fine, int ncomp, const T* f, T* c,
                     hipStream_t s, bool split = false);
template <typename T>
int mg_restrict_side(const Level& coarse, const Level& fine, const T* f, T* c, hipStream_t s);
// order 0: constant refinement, 1: tri/bilinear (coarse ghosts filled).
template <typename T>
int mg_prolong_add(const Level& fine, const Level& coarse, int ncomp, int order, const T* xc,
                   T* xf, hipStream_t s, bool split = false);
// coarse holds the boxes of fine halved, patch by patch (even corners/extents)
bool is_coarsening(const Level& coarse, const Level& fine);
// After red-black sweeps ending with the black colour, the orientation
// residual vanishes on black cells when every black cell's neighbours are
// red: one checkerboard over the patches (non-negative, even patch corners)
// and periodic ghosts; one, two or four components.
bool red_residual_ok(const Level& L, int ncomp);

// Orientation field (ql <= 4 components) coarsened by the normalised volume
// average, or refined by the normalised (bi/tri)linear interpolation (coarse
// ghosts filled); dxf = fine mesh spacing (quat_transfer.hip).
int quat_coarsen(const Level& coarse, const Level& fine, const double* dxf, int ql,
                 const double* qf, double* qc, hipStream_t s);
int linear_refine(const Level& fine, const Level& coarse, const double* dxf, int ncomp,
                  const double* c, double* f, hipStream_t s);
int quat_refine(const Level& fine, const Level& coarse, const double* dxf, int ql,
                const double* qc, double* qf, hipStream_t s);

}  // namespace ampe
