// Grain identification: Grains::findAndNumberGrains (source/Grains.cc:263-645).
//
// A cell belongs to a grain when its phase is at least the threshold; grains
// are the 6-connected (4-connected in 2-D) components of such cells across
// patch and periodic boundaries.  The reference gives every such cell its
// index in the level's bounding box, lowers each cell to a smaller neighbour
// label until no label changes, and renumbers the surviving labels (the
// smallest box index of each grain) 0, 1, ... in increasing order.  The
// result is a property of the components alone, so the GPU reaches the same
// fixed point its own way: every pass takes the minimum over the cell and
// its neighbours and then follows labels through the cells they name (a label
// is a box index, so lab[cell(label)] is again a label of the same grain),
// which shortens the number of passes from the grain diameter to a few; the
// renumbering is an exclusive scan of the root flags in box order.
#include <algorithm>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace ampe {

namespace {

struct BoxIndex {
   int lo[3];
   long long n[3];
};

__device__ __forceinline__ long long box_id(const BoxIndex& b, const CellPos& c)
{
   long long x = c.pd.lo[0] + c.i - b.lo[0];
   long long y = c.pd.lo[1] + c.j - b.lo[1];
   long long z = c.pd.lo[2] + c.k - b.lo[2];
   return x + b.n[0] * (y + b.n[1] * z);
}

__global__ void k_fill_int(int* __restrict__ a, long long n, int v)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t < n) a[t] = v;
}

// lab = box index where phase >= thr (else -1); cell[box index] = storage index
__global__ void __launch_bounds__(256) k_grain_init(LevelArgs a, BoxIndex b,
                                                    const double* __restrict__ phase,
                                                    double thr, int* __restrict__ lab,
                                                    int* __restrict__ cell)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   long long id = box_id(b, c);
   lab[c.c0] = phase[c.c0] >= thr ? (int)id : -1;
   cell[id] = (int)c.c0;
}

// One pass: minimum over the cell and its face neighbours, then up to
// kHops steps along lab[cell[label]].
constexpr int kHops = 16;

template <int NDIM>
__global__ void __launch_bounds__(256) k_grain_pass(LevelArgs a, int* __restrict__ lab,
                                                    const int* __restrict__ cell,
                                                    int* __restrict__ changed)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   const int own = lab[c.c0];
   if (own < 0) return;
   int m = own;
#pragma unroll
   for (int d = 0; d < NDIM; d++) {
      int lm = lab[c.c0 - c.st[d]], lp = lab[c.c0 + c.st[d]];
      if (lm >= 0 && lm < m) m = lm;
      if (lp >= 0 && lp < m) m = lp;
   }
   for (int h = 0; h < kHops; h++) {
      int l = lab[cell[m]];
      if (l < 0 || l >= m) break;
      m = l;
   }
   if (m < own) {
      lab[c.c0] = m;
      *changed = 1;
   }
}

// root[box index] = 1 for the cell holding its grain's smallest box index
__global__ void __launch_bounds__(256) k_grain_roots(LevelArgs a, BoxIndex b,
                                                     const int* __restrict__ lab,
                                                     int* __restrict__ root)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   long long id = box_id(b, c);
   root[id] = (lab[c.c0] == (int)id) ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_grain_out(LevelArgs a, const int* __restrict__ lab,
                                                   const int* __restrict__ rank,
                                                   int* __restrict__ out)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   int l = lab[c.c0];
   out[int_idx(c)] = l >= 0 ? rank[l] : -1;
}

struct DevBuf {
   void* p = nullptr;
   ~DevBuf() { (void)hipFree(p); }
};

}  // namespace

int grains_label(const Level& L, const double* phase, double thr, int* labels, int* ngrains,
                 int* passes, hipStream_t s)
{
   BoxIndex b;
   int hi[3];
   for (int d = 0; d < 3; d++) {
      b.lo[d] = L.hp[0].lo[d];
      hi[d] = L.hp[0].lo[d];
   }
   for (const PatchDev& p : L.hp) {
      int n[3] = {p.nx, p.ny, p.nz};
      for (int d = 0; d < 3; d++) {
         b.lo[d] = std::min(b.lo[d], p.lo[d]);
         hi[d] = std::max(hi[d], p.lo[d] + n[d]);
      }
   }
   long long nbox = 1;
   for (int d = 0; d < 3; d++) {
      b.n[d] = hi[d] - b.lo[d];
      nbox *= b.n[d];
   }
   if (nbox >= (1LL << 31) - 1 || L.ncg >= (1LL << 31)) {
      set_error("grains_label: the level's bounding box must have fewer than 2^31 cells");
      return -1;
   }
   DevBuf lab, cell, root, rank, flag, tmp;
   AMPE_HIP_CHECK(hipMalloc(&lab.p, sizeof(int) * (size_t)L.ncg));
   AMPE_HIP_CHECK(hipMalloc(&cell.p, sizeof(int) * (size_t)nbox));
   AMPE_HIP_CHECK(hipMalloc(&root.p, sizeof(int) * (size_t)(nbox + 1)));
   AMPE_HIP_CHECK(hipMalloc(&rank.p, sizeof(int) * (size_t)(nbox + 1)));
   AMPE_HIP_CHECK(hipMalloc(&flag.p, sizeof(int)));
   int* labp = (int*)lab.p;
   int* cellp = (int*)cell.p;
   int* rootp = (int*)root.p;
   int* rankp = (int*)rank.p;
   int* flagp = (int*)flag.p;
   auto fill = [&](int* a, long long n, int v) -> int {
      hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, v);
      AMPE_HIP_CHECK(hipGetLastError());
      return 0;
   };
   // ghosts without a source (physical boundary) stay -1
   if (fill(labp, L.ncg, -1) || fill(cellp, nbox, 0) || fill(rootp, nbox + 1, 0)) return -1;
   LevelArgs a = level_args(L);
   dim3 g(L.ncell_tiles), blk(TX, TY);
   hipLaunchKernelGGL(k_grain_init, g, blk, 0, s, a, b, phase, thr, labp, cellp);
   AMPE_HIP_CHECK(hipGetLastError());
   // every pass lowers some label or ends the loop (labels only decrease), so
   // the loop ends; the bound only guards against a grain whose cells form a
   // path longer than any realistic microstructure
   const long long maxpass = 100000;
   int np = 0, h = 1;
   while (h) {
      if (np >= maxpass) {
         set_error("grains_label: labels did not settle");
         return -1;
      }
      if (fill(flagp, 1, 0)) return -1;
      if (ghost_fill(L, labp, 1, s)) return -1;
      if (L.ndim == 3)
         hipLaunchKernelGGL(k_grain_pass<3>, g, blk, 0, s, a, labp, cellp, flagp);
      else
         hipLaunchKernelGGL(k_grain_pass<2>, g, blk, 0, s, a, labp, cellp, flagp);
      AMPE_HIP_CHECK(hipGetLastError());
      AMPE_HIP_CHECK(hipMemcpyAsync(L.hpinned, flagp, sizeof(int), hipMemcpyDeviceToHost, s));
      AMPE_HIP_CHECK(hipStreamSynchronize(s));
      std::memcpy(&h, L.hpinned, sizeof(int));
      np++;
   }
   hipLaunchKernelGGL(k_grain_roots, g, blk, 0, s, a, b, labp, rootp);
   AMPE_HIP_CHECK(hipGetLastError());
   size_t tbytes = 0;
   AMPE_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, rootp, rankp,
                                                   (int)(nbox + 1), s));
   AMPE_HIP_CHECK(hipMalloc(&tmp.p, tbytes));
   AMPE_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tbytes, rootp, rankp, (int)(nbox + 1),
                                                   s));
   hipLaunchKernelGGL(k_grain_out, g, blk, 0, s, a, labp, rankp, labels);
   AMPE_HIP_CHECK(hipGetLastError());
   AMPE_HIP_CHECK(hipMemcpyAsync(L.hpinned, rankp + nbox, sizeof(int), hipMemcpyDeviceToHost, s));
   AMPE_HIP_CHECK(hipStreamSynchronize(s));
   int ng = 0;
   std::memcpy(&ng, L.hpinned, sizeof(int));
   if (ngrains) *ngrains = ng;
   if (passes) *passes = np;
   return 0;
}

}  // namespace ampe
