// Level layout, ghost gather and deterministic reductions.
#include "level.h"
#include "kernels.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace ampe {

static std::mutex g_err_mu;
static std::string g_err;

void set_error(const std::string& msg)
{
   std::lock_guard<std::mutex> lk(g_err_mu);
   g_err = msg;
}

const char* last_error()
{
   std::lock_guard<std::mutex> lk(g_err_mu);
   return g_err.c_str();
}

static int wrapi(int i, int lo, int n)
{
   int r = (i - lo) % n;
   if (r < 0) r += n;
   return lo + r;
}

// Dispatch order of a tile list.  Workgroups are dispatched round-robin over
// the 8 XCDs (workgroup b to XCD b % 8), each XCD with its own L2.  Every
// plane (patch, k) of the list is cut into 8 contiguous chunks of rows, chunk
// x of every plane going to XCD x in plane order: an XCD sweeps one band of
// rows through all planes, so the x/y halos of its tiles and the previous
// planes it reads are in its own L2.  The per-XCD sequences are concatenated
// and dealt out so that workgroup b takes the (b / 8)-th tile of XCD b % 8's
// share.
static void xcd_order(std::vector<Tile>& tiles)
{
   constexpr int X = 8;
   const size_t n = tiles.size();
   if (n == 0) return;
   std::vector<std::vector<Tile>> seq(X);
   size_t p0 = 0;
   while (p0 < n) {
      size_t p1 = p0;
      while (p1 < n && tiles[p1].patch == tiles[p0].patch && tiles[p1].k == tiles[p0].k) p1++;
      size_t tp = p1 - p0;
      for (int x = 0; x < X; x++) {
         size_t a = p0 + tp * x / X, b = p0 + tp * (x + 1) / X;
         seq[x].insert(seq[x].end(), tiles.begin() + a, tiles.begin() + b);
      }
      p0 = p1;
   }
   std::vector<Tile> q;
   q.reserve(n);
   for (int x = 0; x < X; x++) q.insert(q.end(), seq[x].begin(), seq[x].end());
   const size_t qq = n / X, r = n % X;
   for (size_t b = 0; b < n; b++) {
      size_t x = b % X, slot = b / X;
      tiles[b] = q[x * qq + (x < r ? x : r) + slot];
   }
}

// Work lists, device copies of the patch table and the ghost gather, and
// the reduction buffers of a level whose patch table L->hp is complete.
static int level_upload(Level* L, const std::vector<long long>& gd,
                        const std::vector<long long>& gs, Level** out)
{
   const int ndim = L->ndim, npatch = L->npatch;
   // work lists
   std::vector<Tile> ct, ft, zt, mt;
   for (int p = 0; p < npatch; p++) {
      const PatchDev& d = L->hp[p];
      for (int k = 0; k < d.nz; k++)
         for (int j0 = 0; j0 < d.ny; j0 += TY)
            for (int i0 = 0; i0 < d.nx; i0 += TX) ct.push_back({p, i0, j0, k});
      for (int k = 0; k < d.nz; k += (ndim == 3 ? KZ : 1))
         for (int j0 = 0; j0 < d.ny; j0 += TY)
            for (int i0 = 0; i0 < d.nx; i0 += TX) zt.push_back({p, i0, j0, k});
      for (int k = 0; k < d.nz; k += (ndim == 3 ? KM : 1))
         for (int j0 = 0; j0 < d.ny; j0 += TY)
            for (int i0 = 0; i0 < d.nx; i0 += TX) mt.push_back({p, i0, j0, k});
      int fz = (ndim == 3) ? d.nz + 1 : 1;
      for (int k = 0; k < fz; k++)
         for (int j0 = 0; j0 < d.ny + 1; j0 += TY)
            for (int i0 = 0; i0 < d.nx + 1; i0 += TX) ft.push_back({p, i0, j0, k});
   }
   xcd_order(ct);
   xcd_order(ft);
   xcd_order(zt);
   xcd_order(mt);
   L->nmarch_tiles = (int)mt.size();
   L->ncell_tiles = (int)ct.size();
   L->nface_tiles = (int)ft.size();
   L->ncol_tiles = (int)zt.size();

   L->nghost = (long long)gd.size();

   int maxparts = std::max(L->ncell_tiles, L->nface_tiles);
   hipError_t e = hipSuccess;
   e = hipMalloc(&L->dp, sizeof(PatchDev) * npatch);
   if (e == hipSuccess) e = hipMalloc(&L->cell_tiles, sizeof(Tile) * ct.size());
   if (e == hipSuccess) e = hipMalloc(&L->face_tiles, sizeof(Tile) * ft.size());
   if (e == hipSuccess) e = hipMalloc(&L->col_tiles, sizeof(Tile) * zt.size());
   if (e == hipSuccess) e = hipMalloc(&L->march_tiles, sizeof(Tile) * mt.size());
   if (e == hipSuccess && L->nghost > 0) {
      e = hipMalloc(&L->ghost_dst, sizeof(long long) * gd.size());
      if (e == hipSuccess) e = hipMalloc(&L->ghost_src, sizeof(long long) * gs.size());
   }
   if (e == hipSuccess) e = hipMalloc(&L->partial, sizeof(double) * (size_t)maxparts * 16);
   if (e == hipSuccess) e = hipMalloc(&L->dred, sizeof(double) * 16);
   if (e == hipSuccess) e = hipHostMalloc(&L->hpinned, sizeof(double) * 64);
   if (e == hipSuccess) e = hipMalloc(&L->iscratch, sizeof(int) * 64);
   if (e == hipSuccess)
      // KKS partials: one slot per 256 ghosted cells plus 256 slots of the
      // surrogate pass
      e = hipMalloc(&L->kks_part,
                    sizeof(double) * 8 * ((size_t)((L->ncg + 255) / 256 + 1) + 256));
   if (e == hipSuccess)
      e = hipMemcpy(L->dp, L->hp.data(), sizeof(PatchDev) * npatch, hipMemcpyHostToDevice);
   if (e == hipSuccess)
      e = hipMemcpy(L->cell_tiles, ct.data(), sizeof(Tile) * ct.size(), hipMemcpyHostToDevice);
   if (e == hipSuccess)
      e = hipMemcpy(L->face_tiles, ft.data(), sizeof(Tile) * ft.size(), hipMemcpyHostToDevice);
   if (e == hipSuccess)
      e = hipMemcpy(L->col_tiles, zt.data(), sizeof(Tile) * zt.size(), hipMemcpyHostToDevice);
   if (e == hipSuccess)
      e = hipMemcpy(L->march_tiles, mt.data(), sizeof(Tile) * mt.size(), hipMemcpyHostToDevice);
   if (e == hipSuccess && L->nghost > 0) {
      e = hipMemcpy(L->ghost_dst, gd.data(), sizeof(long long) * gd.size(), hipMemcpyHostToDevice);
      if (e == hipSuccess)
         e = hipMemcpy(L->ghost_src, gs.data(), sizeof(long long) * gs.size(), hipMemcpyHostToDevice);
   }
   if (e != hipSuccess) {
      set_error(std::string("level_create: ") + hipGetErrorString(e));
      level_destroy(L);
      return -1;
   }
   *out = L;
   return 0;
}


int level_create(int ndim, const ampe_box* boxes, int npatch, int ghost,
                 const ampe_box* periodic, Level** out)
{
   if (ndim != 2 && ndim != 3) {
      set_error("level_create: ndim must be 2 or 3");
      return -1;
   }
   if (npatch <= 0 || !boxes || ghost < 1) {
      set_error("level_create: need npatch > 0 boxes and ghost >= 1");
      return -1;
   }
   Level* L = new Level();
   L->ndim = ndim;
   L->g = ghost;
   L->gz = (ndim == 3) ? ghost : 0;
   L->npatch = npatch;
   if (periodic) {
      L->periodic = true;
      L->pbox = *periodic;
   }
   L->hp.resize(npatch);
   long long coff = 0, ioff = 0, soff[3] = {0, 0, 0};
   for (int p = 0; p < npatch; p++) {
      PatchDev& d = L->hp[p];
      int n[3];
      for (int a = 0; a < 3; a++) {
         n[a] = boxes[p].hi[a] - boxes[p].lo[a] + 1;
         d.lo[a] = boxes[p].lo[a];
      }
      if (ndim == 2) {
         n[2] = 1;
         d.lo[2] = 0;
      }
      if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) {
         set_error("level_create: empty box");
         delete L;
         return -1;
      }
      d.nx = n[0];
      d.ny = n[1];
      d.nz = n[2];
      d.sy = n[0] + 2 * ghost;
      d.sz = (long long)d.sy * (n[1] + 2 * ghost);
      d.cell_off = coff;
      coff += d.sz * (n[2] + 2 * L->gz);
      d.int_off = ioff;
      ioff += (long long)n[0] * n[1] * n[2];
      for (int a = 0; a < 3; a++) {
         d.side_off[a] = soff[a];
         if (a < ndim)
            soff[a] += (long long)(n[0] + (a == 0)) * (n[1] + (a == 1)) *
                       (n[2] + (a == 2));
      }
   }
   // sides on the physical boundary (SAMRAI's getTouchesRegularBoundary): the
   // level's bounding box stands for the domain, as every level built here
   // covers it; none on a periodic domain
   int blo[3], bhi[3];  // bounding box, bhi exclusive
   {
      for (int a = 0; a < 3; a++) {
         blo[a] = L->hp[0].lo[a];
         bhi[a] = blo[a];
      }
      for (const PatchDev& d : L->hp) {
         int n[3] = {d.nx, d.ny, d.nz};
         for (int a = 0; a < 3; a++) {
            blo[a] = std::min(blo[a], d.lo[a]);
            bhi[a] = std::max(bhi[a], d.lo[a] + n[a]);
         }
      }
      for (PatchDev& d : L->hp) {
         int n[3] = {d.nx, d.ny, d.nz};
         d.physb = 0;
         if (periodic) continue;
         for (int a = 0; a < ndim; a++) {
            if (d.lo[a] == blo[a]) d.physb |= 1 << (2 * a);
            if (d.lo[a] + n[a] == bhi[a]) d.physb |= 1 << (2 * a + 1);
         }
      }
   }
   L->ncg = coff;
   L->ncell = ioff;
   for (int a = 0; a < 3; a++) L->nside[a] = soff[a];
   L->side_dir_off[0] = 0;
   L->side_dir_off[1] = soff[0];
   L->side_dir_off[2] = soff[0] + soff[1];
   L->nside_tot = soff[0] + soff[1] + soff[2];

   // ghost gather: each ghost cell copies the interior cell it images
   std::vector<long long> gd, gs;
   for (int p = 0; p < npatch; p++) {
      const PatchDev& d = L->hp[p];
      int g = ghost, gz = L->gz;
      for (int k = -gz; k < d.nz + gz; k++)
         for (int j = -g; j < d.ny + g; j++)
            for (int i = -g; i < d.nx + g; i++) {
               if (i >= 0 && i < d.nx && j >= 0 && j < d.ny && k >= 0 && k < d.nz)
                  continue;
               int gi[3] = {d.lo[0] + i, d.lo[1] + j, d.lo[2] + k};
               if (periodic) {
                  for (int a = 0; a < ndim; a++)
                     gi[a] = wrapi(gi[a], periodic->lo[a],
                                   periodic->hi[a] - periodic->lo[a] + 1);
               } else {
                  // physical boundary: zero-slope image inside the bounding
                  // box, then whichever patch owns it, so a ghost takes the
                  // same value however the domain is cut into patches
                  for (int a = 0; a < ndim; a++)
                     gi[a] = std::min(std::max(gi[a], blo[a]), bhi[a] - 1);
               }
               int src_p = -1, si = 0, sj = 0, sk = 0;
               for (int q = 0; q < npatch && src_p < 0; q++) {
                  const PatchDev& e = L->hp[q];
                  int li = gi[0] - e.lo[0], lj = gi[1] - e.lo[1],
                      lk = (ndim == 3) ? gi[2] - e.lo[2] : 0;
                  if (li >= 0 && li < e.nx && lj >= 0 && lj < e.ny && lk >= 0 &&
                      lk < e.nz) {
                     src_p = q;
                     si = li;
                     sj = lj;
                     sk = lk;
                  }
               }
               if (src_p < 0) {
                  // image not covered by any patch (a level that does not
                  // fill its bounding box): zero-slope extension of this patch
                  src_p = p;
                  si = std::min(std::max(i, 0), d.nx - 1);
                  sj = std::min(std::max(j, 0), d.ny - 1);
                  sk = std::min(std::max(k, 0), d.nz - 1);
               }
               const PatchDev& e = L->hp[src_p];
               gd.push_back(d.cell_off + (i + g) + (long long)(j + g) * d.sy +
                            (long long)(k + gz) * d.sz);
               gs.push_back(e.cell_off + (si + g) + (long long)(sj + g) * e.sy +
                            (long long)(sk + gz) * e.sz);
            }
   }
   return level_upload(L, gd, gs, out);
}

int level_create_placed(int ndim, const ampe_box* boxes, int npatch, const LevelPlacement& pl,
                        const std::vector<long long>& gd, const std::vector<long long>& gs,
                        Level** out)
{
   if ((ndim != 2 && ndim != 3) || npatch <= 0 || !boxes) {
      set_error("level_create_placed: bad arguments");
      return -1;
   }
   Level* L = new Level();
   L->ndim = ndim;
   L->g = 1;
   L->gz = (ndim == 3) ? 1 : 0;
   L->npatch = npatch;
   L->hp.resize(npatch);
   long long coff = pl.cell_base, ioff = pl.int_base;
   long long soff[3] = {pl.side_base[0], pl.side_base[1], pl.side_base[2]};
   for (int p = 0; p < npatch; p++) {
      PatchDev& d = L->hp[p];
      int n[3];
      for (int a = 0; a < 3; a++) {
         n[a] = boxes[p].hi[a] - boxes[p].lo[a] + 1;
         d.lo[a] = boxes[p].lo[a];
      }
      if (ndim == 2) {
         n[2] = 1;
         d.lo[2] = 0;
      }
      if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) {
         set_error("level_create_placed: empty box");
         delete L;
         return -1;
      }
      d.nx = n[0];
      d.ny = n[1];
      d.nz = n[2];
      d.sy = n[0] + 2;
      d.sz = (long long)d.sy * (n[1] + 2);
      d.cell_off = coff;
      coff += d.sz * (n[2] + 2 * L->gz);
      d.int_off = ioff;
      ioff += (long long)n[0] * n[1] * n[2];
      d.physb = 0;
      for (int a = 0; a < 3; a++) {
         d.side_off[a] = soff[a];
         if (a < ndim)
            soff[a] += (long long)(n[0] + (a == 0)) * (n[1] + (a == 1)) * (n[2] + (a == 2));
      }
   }
   L->ncg = pl.ncg_total;
   L->ncell = pl.ncell_total;
   for (int a = 0; a < 3; a++) L->nside[a] = pl.nside_total[a];
   L->side_dir_off[0] = 0;
   L->side_dir_off[1] = pl.nside_total[0];
   L->side_dir_off[2] = pl.nside_total[0] + pl.nside_total[1];
   L->nside_tot = pl.nside_total[0] + pl.nside_total[1] + pl.nside_total[2];
   if (coff > pl.ncg_total || ioff > pl.ncell_total) {
      set_error("level_create_placed: patches exceed the composite field");
      delete L;
      return -1;
   }
   return level_upload(L, gd, gs, out);
}

void level_destroy(Level* L)
{
   if (!L) return;
   (void)hipFree(L->dp);
   (void)hipFree(L->cell_tiles);
   (void)hipFree(L->face_tiles);
   (void)hipFree(L->col_tiles);
   (void)hipFree(L->march_tiles);
   (void)hipFree(L->ghost_dst);
   (void)hipFree(L->ghost_src);
   (void)hipFree(L->ghost_dst_s);
   (void)hipFree(L->ghost_src_s);
   (void)hipFree(L->partial);
   (void)hipFree(L->dred);
   (void)hipHostFree(L->hpinned);
   (void)hipFree(L->iscratch);
   (void)hipFree(L->kks_part);
   delete L;
}

template <typename T>
__global__ void __launch_bounds__(256) k_ghost_gather(T* __restrict__ f, long long ncg,
                                                      const long long* __restrict__ dst,
                                                      const long long* __restrict__ src,
                                                      long long n)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   long long off = (long long)blockIdx.y * ncg;
   f[off + dst[t]] = f[off + src[t]];
}

template <typename T>
static int ghost_fill_t(const Level& L, T* f, int ncomp, hipStream_t s, bool split)
{
   if (L.nghost == 0 || ncomp <= 0) return 0;
   if (split && !L.ghost_dst_s) {
      set_error("ghost_fill: the level has no split-layout gather");
      return -1;
   }
   dim3 grid((unsigned)((L.nghost + 255) / 256), (unsigned)ncomp);
   hipLaunchKernelGGL(k_ghost_gather<T>, grid, dim3(256), 0, s, f, L.ncg,
                      split ? L.ghost_dst_s : L.ghost_dst, split ? L.ghost_src_s : L.ghost_src,
                      L.nghost);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int ghost_fill(const Level& L, double* f, int ncomp, hipStream_t s, bool split)
{
   return ghost_fill_t(L, f, ncomp, s, split);
}
int ghost_fill(const Level& L, float* f, int ncomp, hipStream_t s, bool split)
{
   return ghost_fill_t(L, f, ncomp, s, split);
}
int ghost_fill(const Level& L, int* f, int ncomp, hipStream_t s, bool split)
{
   return ghost_fill_t(L, f, ncomp, s, split);
}

__global__ void __launch_bounds__(256) k_split_map(const PatchDev* __restrict__ P, int np,
                                                   const long long* __restrict__ in,
                                                   long long* __restrict__ out, long long n)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t < n) out[t] = split_index(P, np, in[t]);
}

int level_split_tables(Level& L, hipStream_t s)
{
   if (L.ghost_dst_s || L.nghost == 0) return 0;
   for (int p = 1; p < L.npatch; p++)
      if (L.hp[p].cell_off <= L.hp[p - 1].cell_off) {
         set_error("level_split_tables: patch offsets are not ascending");
         return -1;
      }
   size_t bytes = sizeof(long long) * (size_t)L.nghost;
   AMPE_HIP_CHECK(hipMalloc(&L.ghost_dst_s, bytes));
   AMPE_HIP_CHECK(hipMalloc(&L.ghost_src_s, bytes));
   dim3 grid((unsigned)((L.nghost + 255) / 256));
   hipLaunchKernelGGL(k_split_map, grid, dim3(256), 0, s, L.dp, L.npatch, L.ghost_dst,
                      L.ghost_dst_s, L.nghost);
   hipLaunchKernelGGL(k_split_map, grid, dim3(256), 0, s, L.dp, L.npatch, L.ghost_src,
                      L.ghost_src_s, L.nghost);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

// Single-workgroup fixed-order reduction of nparts x ncol partials
// (column-interleaved).  Each lane walks a fixed stride, then a fixed tree.
__global__ void __launch_bounds__(1024) k_reduce_partials(const double* __restrict__ part,
                                                          int nparts, int ncol, int op,
                                                          double* __restrict__ out)
{
   __shared__ double sh[1024];
   for (int c = 0; c < ncol; c++) {
      double acc = 0.0;  // sums start at 0; maxima are of non-negative values
      for (int i = threadIdx.x; i < nparts; i += 1024) {
         double v = part[(long long)i * ncol + c];
         acc = (op == 0) ? acc + v : (v > acc ? v : acc);
      }
      sh[threadIdx.x] = acc;
      __syncthreads();
      for (int w = 512; w > 0; w >>= 1) {
         if (threadIdx.x < w) {
            double a = sh[threadIdx.x], b = sh[threadIdx.x + w];
            sh[threadIdx.x] = (op == 0) ? a + b : (b > a ? b : a);
         }
         __syncthreads();
      }
      if (threadIdx.x == 0) out[c] = sh[0];
      __syncthreads();
   }
}

// One workgroup per region r: the sum of part[r * nparts .. + nparts) in the
// fixed order of k_reduce_partials.
__global__ void __launch_bounds__(1024) k_reduce_regions(const double* __restrict__ part,
                                                         int nparts, double* __restrict__ out)
{
   __shared__ double sh[1024];
   const double* p = part + (long long)blockIdx.x * nparts;
   double acc = 0.0;
   for (int i = threadIdx.x; i < nparts; i += 1024) acc = acc + p[i];
   sh[threadIdx.x] = acc;
   __syncthreads();
   for (int w = 512; w > 0; w >>= 1) {
      if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
      __syncthreads();
   }
   if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

int reduce_partials(const Level& L, int nparts, int ncol, int op,
                    double* host_out, hipStream_t s)
{
   double* dout = L.dred;
   hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, L.partial, nparts,
                      ncol, op, dout);
   AMPE_HIP_CHECK(hipGetLastError());
   AMPE_HIP_CHECK(hipMemcpyAsync(L.hpinned, dout, sizeof(double) * ncol,
                                 hipMemcpyDeviceToHost, s));
   AMPE_HIP_CHECK(hipStreamSynchronize(s));
   for (int c = 0; c < ncol; c++) host_out[c] = L.hpinned[c];
   return 0;
}

int reduce_regions(const Level& L, int nparts, int nreg, double* host_out, hipStream_t s)
{
   if (nreg > 16) {
      set_error("reduce_regions: at most 16 regions");
      return -1;
   }
   hipLaunchKernelGGL(k_reduce_regions, dim3(nreg), dim3(1024), 0, s, L.partial, nparts,
                      L.dred);
   AMPE_HIP_CHECK(hipGetLastError());
   AMPE_HIP_CHECK(hipMemcpyAsync(L.hpinned, L.dred, sizeof(double) * nreg,
                                 hipMemcpyDeviceToHost, s));
   AMPE_HIP_CHECK(hipStreamSynchronize(s));
   for (int r = 0; r < nreg; r++) host_out[r] = L.hpinned[r];
   return 0;
}

// ------------------------------------------------------------ kernel timer
namespace {
struct TimerRec {
   int id;
   hipEvent_t a, b;
   double bytes;
};
bool g_timer_on[KernelTimer::COUNT] = {};
std::vector<TimerRec> g_timer_recs;
std::vector<hipEvent_t> g_timer_pool;

hipEvent_t timer_event()
{
   if (!g_timer_pool.empty()) {
      hipEvent_t e = g_timer_pool.back();
      g_timer_pool.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   if (hipEventCreate(&e) != hipSuccess) return nullptr;
   return e;
}
}  // namespace

KernelTimer::Span::Span(int id, hipStream_t s, double bytes)
{
   if (id < 0 || id >= COUNT || !g_timer_on[id]) return;
   // a launch recorded into a HIP graph is not timed (its events would never
   // complete as such)
   hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
   if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
   TimerRec r{id, timer_event(), timer_event(), bytes};
   if (!r.a || !r.b || hipEventRecord(r.a, s) != hipSuccess) {
      if (r.a) g_timer_pool.push_back(r.a);
      if (r.b) g_timer_pool.push_back(r.b);
      return;
   }
   g_timer_recs.push_back(r);
   slot = (int)g_timer_recs.size() - 1;
   stream = s;
}

KernelTimer::Span::~Span()
{
   if (slot >= 0) (void)hipEventRecord(g_timer_recs[slot].b, stream);
}

int KernelTimer::enable(int id, bool on)
{
   if (id < 0 || id >= COUNT) {
      set_error("kernel timer: unknown kernel id");
      return -1;
   }
   g_timer_on[id] = on;
   return 0;
}

int KernelTimer::read(int id, double* ms, long long* launches, double* bytes)
{
   if (id < 0 || id >= COUNT) {
      set_error("kernel timer: unknown kernel id");
      return -1;
   }
   double tot = 0.0, by = 0.0;
   long long n = 0;
   std::vector<TimerRec> keep;
   for (const TimerRec& r : g_timer_recs) {
      if (r.id != id) {
         keep.push_back(r);
         continue;
      }
      float t = 0.0f;
      AMPE_HIP_CHECK(hipEventSynchronize(r.b));
      AMPE_HIP_CHECK(hipEventElapsedTime(&t, r.a, r.b));
      tot += t;
      by += r.bytes;
      n++;
      g_timer_pool.push_back(r.a);
      g_timer_pool.push_back(r.b);
   }
   g_timer_recs.swap(keep);
   if (ms) *ms = tot;
   if (launches) *launches = n;
   if (bytes) *bytes = by;
   return 0;
}

double quat_rbgs_fused_bytes(const Level& L, int ncomp, int elem)
{
   double n = (double)L.ncell;
   return (double)elem * (3.0 * ncomp * n + (double)L.nside_tot + n);
}

double quat_rbgs_bytes(const Level& L, int ncomp, int elem)
{
   double n = (double)L.ncell;
   return (double)elem * (ncomp * n            // z, all cells
                          + ncomp * 0.5 * n    // rhs, this colour
                          + ncomp * 0.5 * n    // z written, this colour
                          + (double)L.nside_tot  // face coefficients
                          + n);                // sqrt(M)
}

}  // namespace ampe
