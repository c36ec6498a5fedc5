// Patch-level layout on the device: uneven boxes stored back to back, each a
// ghosted cell block (i fastest), side fields per direction without ghosts,
// and the work lists every kernel launches over (one launch per level).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ampe_hip.h"

namespace ampe {

void set_error(const std::string& msg);

#define AMPE_HIP_CHECK(call)                                              \
   do {                                                                   \
      hipError_t e_ = (call);                                             \
      if (e_ != hipSuccess) {                                             \
         ::ampe::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
         return -1;                                                       \
      }                                                                   \
   } while (0)

// Tile of cells processed by one workgroup: TX x TY cells of one (patch, k).
constexpr int TX = 64;
constexpr int TY = 4;
// Planes per column tile: a workgroup of a column kernel walks KZ consecutive
// k planes, so the k-neighbour loads of a plane are the previous plane's
// centre loads (cache hits) instead of HBM reads.
constexpr int KZ = 1;
// Planes per column of the marching kernels (the one-pass red-black sweep):
// a workgroup walks KM planes keeping three planes of results in LDS.
constexpr int KM = 16;

struct PatchDev {
   int nx, ny, nz;         // interior extents (nz = 1 in 2-D)
   int sy;                 // ghost-box row length nx + 2g
   long long sz;           // ghost-box plane size
   long long cell_off;     // ghost-box offset inside one component block
   long long side_off[3];  // patch faces offset inside direction block d
   long long int_off;      // compact interior offset (host I/O, tags)
   int lo[3];              // box lower corner (global index space)
   int physb;              // bit 2d / 2d+1: lower / upper side in direction d lies on a
                           // physical (non-periodic) boundary of the level's bounding box
};

struct Tile {
   int patch, i0, j0, k;
};

struct Level {
   int ndim = 2, g = 1, gz = 0, npatch = 0;
   bool periodic = false;       // periodic domain (pbox) or physical boundaries
   ampe_box pbox = {{0, 0, 0}, {0, 0, 0}};
   long long ncg = 0;         // ghosted cells per component
   long long ncell = 0;       // interior cells
   long long nside[3] = {0, 0, 0};
   long long side_dir_off[3] = {0, 0, 0};
   long long nside_tot = 0;
   std::vector<PatchDev> hp;
   PatchDev* dp = nullptr;
   Tile* cell_tiles = nullptr;  // interior cells
   int ncell_tiles = 0;
   Tile* face_tiles = nullptr;  // cells grown by one in each direction
   int nface_tiles = 0;
   Tile* col_tiles = nullptr;   // interior cells in columns of KZ planes (3-D)
   int ncol_tiles = 0;
   Tile* march_tiles = nullptr;  // interior cells in columns of KM planes (3-D)
   int nmarch_tiles = 0;
   long long* ghost_dst = nullptr;  // ghost gather (periodic / neighbours)
   long long* ghost_src = nullptr;
   long long nghost = 0;
   // the same gather in the row parity-split layout (built on request by
   // level_split_tables)
   long long* ghost_dst_s = nullptr;
   long long* ghost_src_s = nullptr;
   double* partial = nullptr;  // per-workgroup reduction partials (16 per tile)
   double* dred = nullptr;     // reduction results (16 doubles)
   double* hpinned = nullptr;  // pinned host scalars
   int* iscratch = nullptr;    // device integer flags
   double* kks_part = nullptr; // KKS per-workgroup partials (8 per 256 ghosted cells)
   // control-volume weights of the cells in dot products and norms (a
   // hierarchy's composite level: zero under a finer level); null: 1
   const double* wt = nullptr;
};

int level_create(int ndim, const ampe_box* boxes, int npatch, int ghost,
                 const ampe_box* periodic, Level** out);
void level_destroy(Level* L);

// One level of a hierarchy placed in composite fields: its patches' ghosted
// blocks from cell_base inside each component block of ncg_total cells, its
// faces from side_base[d] inside direction block d of a side field with
// nside_total[d] faces per direction, interior (compact) offsets from
// int_base; ghost width 1; gd / gs: its ghost gather (composite indices).
struct LevelPlacement {
   long long cell_base = 0, ncg_total = 0, int_base = 0, ncell_total = 0;
   long long side_base[3] = {0, 0, 0}, nside_total[3] = {0, 0, 0};
};
int level_create_placed(int ndim, const ampe_box* boxes, int npatch, const LevelPlacement& pl,
                        const std::vector<long long>& gd, const std::vector<long long>& gs,
                        Level** out);

// Ghost gather for ncomp component blocks (double or float fields); split:
// the blocks are stored in the row parity-split layout (split_col), which
// needs level_split_tables first.
int ghost_fill(const Level& L, double* f, int ncomp, hipStream_t s, bool split = false);
int ghost_fill(const Level& L, float* f, int ncomp, hipStream_t s, bool split = false);
int ghost_fill(const Level& L, int* f, int ncomp, hipStream_t s, bool split = false);
// Builds the split-layout ghost gather of L (once; later calls return 0).
int level_split_tables(Level& L, hipStream_t s);

// Fixed-order device reduction of the per-tile partials written by a cell
// kernel: sum (op = 0) or max (op = 1) of nparts values (ncol columns
// interleaved), result to host.
int reduce_partials(const Level& L, int nparts, int ncol, int op,
                    double* host_out, hipStream_t s);
// Sums of nreg consecutive regions of nparts partials each (region r starts at
// L.partial + r * nparts), one launch and one synchronisation.
int reduce_regions(const Level& L, int nparts, int nreg, double* host_out, hipStream_t s);

// Optional event timing of selected kernels, process-wide.  While a kernel
// class is enabled, each of its launches is bracketed by two events recorded
// on its launch stream, with the launch's algorithmic bytes; read() waits for
// the recorded launches and returns their total duration and bytes.  Not
// thread-safe; launches under hipGraph capture must not be timed.
struct KernelTimer {
   enum Id { QUAT_RBGS = 0, COUNT = 1 };
   struct Span {
      Span(int id, hipStream_t s, double bytes);
      ~Span();
      int slot = -1;
      hipStream_t stream = nullptr;
   };
   static int enable(int id, bool on);
   static int read(int id, double* ms, long long* launches, double* bytes);
};

// Algorithmic bytes of one colour of the orientation red-black sweep: all z
// read, the colour's z written and rhs read, every face coefficient and the
// cell sqrt(M) read once.
double quat_rbgs_bytes(const Level& L, int ncomp, int elem);
// The same for one fused sweep (both colours): z read and written, rhs, face
// coefficients and sqrt(M) read once.
double quat_rbgs_fused_bytes(const Level& L, int ncomp, int elem);

}  // namespace ampe
