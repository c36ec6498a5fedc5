// Clustering of tagged cells into boxes: the box-generator role SAMRAI's
// BergerRigoutsos plays in GriddingAlgorithm::regridAllFinerLevels
// (source/PFModel.cc:565-590, mesh::BergerRigoutsos constructed at
// source/PFModel.cc:258), after the tag buffering of bufferTagsOnLevel
// (tag_buffer, source/PFModel.cc:336, source/QuatModel.cc:642-655).
//
// The algorithm is Berger & Rigoutsos (IEEE Trans. Syst. Man Cybern. 21
// (1991) 1278-1286) restated as below; SAMRAI's own asynchronous variant is
// not in this image, so its boxes are not reproduced.  A candidate box is
// first shrunk to the bounding box of its tags; it is accepted when tagged
// cells / cells >= efficiency_tol or when no direction can be cut into two
// pieces of at least min_size cells.  Otherwise it is cut
//   1. at the empty slab (zero signature) nearest its middle, over all
//      directions (relative distance, then lower direction);
//   2. else at the strongest sign change of the signature's second difference
//      (|L[i+1] - L[i]|, ties: nearer the middle, then lower direction);
//   3. else in the middle of its longest cuttable direction;
// every cut leaving at least min_size cells on each side, and both pieces
// are clustered again.  Accepted boxes larger than max_size in a direction
// are then chopped into near-equal pieces.  The tags are first grown by
// tag_buffer cells in every direction (box neighbourhood) and clipped to the
// domain.  Host code: the tags are copied from the device once.
#include <algorithm>
#include <cstring>
#include <vector>

#include "kernels.h"

namespace ampe {

namespace {

struct Mask {
   int lo[3], n[3];
   std::vector<unsigned char> m;
   unsigned char at(int i, int j, int k) const
   {
      return m[(size_t)(i - lo[0]) + (size_t)n[0] * ((size_t)(j - lo[1]) + (size_t)n[1] * (k - lo[2]))];
   }
};

struct Box3 {
   int lo[3], hi[3];  // inclusive
   long long cells() const
   {
      return (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
   }
   int ext(int d) const { return hi[d] - lo[d] + 1; }
};

struct Clusterer {
   const Mask& M;
   int ndim;
   double eff;
   int minsz[3];
   std::vector<Box3> out;

   // signatures of B: sig[d][i] = tagged cells in slab lo[d] + i
   long long signatures(const Box3& B, std::vector<long long> sig[3]) const
   {
      for (int d = 0; d < 3; d++) sig[d].assign((size_t)B.ext(d), 0);
      long long n = 0;
      for (int k = B.lo[2]; k <= B.hi[2]; k++)
         for (int j = B.lo[1]; j <= B.hi[1]; j++)
            for (int i = B.lo[0]; i <= B.hi[0]; i++)
               if (M.at(i, j, k)) {
                  sig[0][i - B.lo[0]]++;
                  sig[1][j - B.lo[1]]++;
                  sig[2][k - B.lo[2]]++;
                  n++;
               }
      return n;
   }

   void run(Box3 B)
   {
      std::vector<long long> sig[3];
      long long n = signatures(B, sig);
      if (n == 0) return;
      // shrink to the tags' bounding box
      for (int d = 0; d < ndim; d++) {
         int a = 0, b = B.ext(d) - 1;
         while (sig[d][a] == 0) a++;
         while (sig[d][b] == 0) b--;
         int lo = B.lo[d];
         B.lo[d] = lo + a;
         B.hi[d] = lo + b;
      }
      n = signatures(B, sig);
      if ((double)n >= eff * (double)B.cells()) {
         out.push_back(B);
         return;
      }
      int cut_d = -1, cut_c = -1;  // cut before slab c: [0, c) | [c, ext)
      // 1. empty slabs
      double best = 1e300;
      for (int d = 0; d < ndim; d++) {
         int e = B.ext(d);
         if (e < 2 * minsz[d]) continue;
         for (int c = minsz[d]; c <= e - minsz[d]; c++) {
            if (sig[d][c] != 0 && sig[d][c - 1] != 0) continue;
            // a hole at slab c (or c - 1): cut between them
            double dist = std::abs((double)c - 0.5 * e) / e;
            if (dist < best) {
               best = dist;
               cut_d = d;
               cut_c = c;
            }
         }
      }
      // 2. inflection of the second difference
      if (cut_d < 0) {
         long long bests = 0;
         double bestd = 1e300;
         for (int d = 0; d < ndim; d++) {
            int e = B.ext(d);
            if (e < 2 * minsz[d] || e < 4) continue;
            std::vector<long long> L((size_t)e, 0);
            for (int i = 1; i + 1 < e; i++) L[i] = sig[d][i - 1] - 2 * sig[d][i] + sig[d][i + 1];
            for (int i = 1; i + 2 < e; i++) {
               if (!((L[i] < 0 && L[i + 1] > 0) || (L[i] > 0 && L[i + 1] < 0))) continue;
               int c = i + 1;
               if (c < minsz[d] || c > e - minsz[d]) continue;
               long long s = L[i + 1] > L[i] ? L[i + 1] - L[i] : L[i] - L[i + 1];
               double dist = std::abs((double)c - 0.5 * e) / e;
               if (s > bests || (s == bests && dist < bestd)) {
                  bests = s;
                  bestd = dist;
                  cut_d = d;
                  cut_c = c;
               }
            }
         }
      }
      // 3. bisection of the longest cuttable direction
      if (cut_d < 0) {
         int bl = 0;
         for (int d = 0; d < ndim; d++)
            if (B.ext(d) >= 2 * minsz[d] && B.ext(d) > bl) {
               bl = B.ext(d);
               cut_d = d;
            }
         if (cut_d < 0) {
            out.push_back(B);
            return;
         }
         cut_c = B.ext(cut_d) / 2;
      }
      Box3 L1 = B, R1 = B;
      L1.hi[cut_d] = B.lo[cut_d] + cut_c - 1;
      R1.lo[cut_d] = B.lo[cut_d] + cut_c;
      run(L1);
      run(R1);
   }
};

void chop(const Box3& B, int ndim, const int* maxsz, std::vector<Box3>& out)
{
   int np[3] = {1, 1, 1};
   for (int d = 0; d < ndim; d++)
      if (maxsz[d] > 0) np[d] = (B.ext(d) + maxsz[d] - 1) / maxsz[d];
   for (int c = 0; c < np[2]; c++)
      for (int b = 0; b < np[1]; b++)
         for (int a = 0; a < np[0]; a++) {
            int idx[3] = {a, b, c};
            Box3 P = B;
            for (int d = 0; d < 3; d++) {
               long long e = B.ext(d);
               P.lo[d] = B.lo[d] + (int)(e * idx[d] / np[d]);
               P.hi[d] = B.lo[d] + (int)(e * (idx[d] + 1) / np[d]) - 1;
            }
            out.push_back(P);
         }
}

}  // namespace

int cluster_mask(int ndim, const ampe_box& dom, const unsigned char* mask,
                 const ampe_cluster_params& cp, ampe_box* boxes, int max_boxes, int* nboxes)
{
   if (ndim != 2 && ndim != 3) {
      set_error("cluster: ndim must be 2 or 3");
      return -1;
   }
   if (!(cp.efficiency_tol > 0.0 && cp.efficiency_tol <= 1.0) || cp.tag_buffer < 0) {
      set_error("cluster: need 0 < efficiency_tol <= 1 and tag_buffer >= 0");
      return -1;
   }
   Mask M;
   for (int d = 0; d < 3; d++) {
      M.lo[d] = d < ndim ? dom.lo[d] : 0;
      M.n[d] = d < ndim ? dom.hi[d] - dom.lo[d] + 1 : 1;
      if (M.n[d] <= 0) {
         set_error("cluster: empty domain");
         return -1;
      }
   }
   const size_t ncell = (size_t)M.n[0] * M.n[1] * M.n[2];
   M.m.assign(mask, mask + ncell);
   // tag buffer: box neighbourhood of tag_buffer cells, clipped to the domain
   if (cp.tag_buffer > 0) {
      std::vector<unsigned char> g = M.m;
      for (int d = 0; d < ndim; d++) {
         std::vector<unsigned char> h(ncell, 0);
         size_t st = d == 0 ? 1 : d == 1 ? (size_t)M.n[0] : (size_t)M.n[0] * M.n[1];
         for (size_t c = 0; c < ncell; c++) {
            if (!g[c]) continue;
            long long x = (long long)((c / st) % (size_t)M.n[d]);
            long long a = std::max<long long>(0, x - cp.tag_buffer);
            long long b = std::min<long long>(M.n[d] - 1, x + cp.tag_buffer);
            for (long long y = a; y <= b; y++) h[c + (size_t)((y - x) * (long long)st)] = 1;
         }
         g.swap(h);
      }
      M.m.swap(g);
   }
   Clusterer C{M, ndim, cp.efficiency_tol, {1, 1, 1}, {}};
   for (int d = 0; d < ndim; d++) C.minsz[d] = std::max(1, cp.min_size[d]);
   Box3 root;
   for (int d = 0; d < 3; d++) {
      root.lo[d] = M.lo[d];
      root.hi[d] = M.lo[d] + M.n[d] - 1;
   }
   C.run(root);
   std::vector<Box3> res;
   for (const Box3& B : C.out) chop(B, ndim, cp.max_size, res);
   if ((int)res.size() > max_boxes) {
      set_error("cluster: more boxes than max_boxes");
      *nboxes = (int)res.size();
      return -1;
   }
   for (size_t b = 0; b < res.size(); b++) {
      for (int d = 0; d < 3; d++) {
         boxes[b].lo[d] = d < ndim ? res[b].lo[d] : 0;
         boxes[b].hi[d] = d < ndim ? res[b].hi[d] : 0;
      }
   }
   *nboxes = (int)res.size();
   return 0;
}

int cluster_tags(const Level& L, const int* tags, const ampe_cluster_params& cp,
                 ampe_box* boxes, int max_boxes, int* nboxes, hipStream_t s)
{
   ampe_box dom;
   for (int d = 0; d < 3; d++) {
      dom.lo[d] = L.hp[0].lo[d];
      dom.hi[d] = L.hp[0].lo[d];
   }
   for (const PatchDev& p : L.hp) {
      int n[3] = {p.nx, p.ny, p.nz};
      for (int d = 0; d < 3; d++) {
         dom.lo[d] = std::min(dom.lo[d], p.lo[d]);
         dom.hi[d] = std::max(dom.hi[d], p.lo[d] + n[d] - 1);
      }
   }
   std::vector<int> h((size_t)L.ncell);
   AMPE_HIP_CHECK(hipMemcpyAsync(h.data(), tags, sizeof(int) * (size_t)L.ncell,
                                 hipMemcpyDeviceToHost, s));
   AMPE_HIP_CHECK(hipStreamSynchronize(s));
   long long nx = dom.hi[0] - dom.lo[0] + 1, ny = dom.hi[1] - dom.lo[1] + 1,
             nz = dom.hi[2] - dom.lo[2] + 1;
   std::vector<unsigned char> mask((size_t)(nx * ny * nz), 0);
   for (const PatchDev& p : L.hp)
      for (int k = 0; k < p.nz; k++)
         for (int j = 0; j < p.ny; j++)
            for (int i = 0; i < p.nx; i++) {
               long long c = p.int_off + i + (long long)p.nx * (j + (long long)p.ny * k);
               if (!h[(size_t)c]) continue;
               long long x = p.lo[0] + i - dom.lo[0], y = p.lo[1] + j - dom.lo[1],
                         z = p.lo[2] + k - dom.lo[2];
               mask[(size_t)(x + nx * (y + ny * z))] = 1;
            }
   return cluster_mask(L.ndim, dom, mask.data(), cp, boxes, max_boxes, nboxes);
}

}  // namespace ampe
