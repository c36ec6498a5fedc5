// Fused vector kernels of the Newton-Krylov loop.  Each runs over the interior
// cells of a level (tile list) and all nc component blocks, performs the
// element-wise update of one or two steps of SPGMR / Newton in one pass and
// writes one partial sum per tile for the reduction that follows, so a
// Gram-Schmidt step costs one read of each operand instead of two or three.
// The element-wise arithmetic is that of the unfused sequence (same order,
// no contraction).  Ghost entries of Krylov vectors are not touched: every
// consumer that reads a stencil ghost-fills its input first.
#include "krylov.h"

namespace ampe {

__device__ __forceinline__ void tile_store_sum(double acc, double* __restrict__ partial)
{
   for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
   __shared__ double sh[4];
   int t = threadIdx.x + TX * threadIdx.y;
   if ((t & 63) == 0) sh[t >> 6] = acc;
   __syncthreads();
   if (t == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// V = hinv V; xk = V / ewt; partial = sum (xk ewt)^2
__global__ void __launch_bounds__(256) k_gmres_prep(LevelArgs a, int nc, double hinv,
                                                    double* __restrict__ V,
                                                    const double* __restrict__ ewt,
                                                    double* __restrict__ xk,
                                                    double* __restrict__ partial)
{
   CellPos c;
   double acc = 0.0;
   if (cell_pos(a, 0, c)) {
      const double w = red_w(a, c.c0);
      for (int m = 0; m < nc; m++) {
         long long i = m * a.ncg + c.c0;
         double v = V[i] * hinv;
         V[i] = v;
         double x = v / ewt[i];
         xk[i] = x;
         double t = x * ewt[i];
         acc += (t * t) * w;
      }
   }
   tile_store_sum(acc, partial);
}

// Krylov direction x = V / ewt: the stored xk, or (V h) / ewt from an
// unnormalised basis vector V with scale h (xk == nullptr).
struct KryDir {
   const double* xk;
   const double* V;
   double h;
   const double* ewt;
};

__device__ __forceinline__ double kry_dir(const KryDir& d, long long i)
{
   return d.xk ? d.xk[i] : (d.V[i] * d.h) / d.ewt[i];
}

// work2 = sig x + ycur
__global__ void __launch_bounds__(256) k_gmres_mid(LevelArgs a, int nc, double sig, KryDir x,
                                                   const double* __restrict__ ycur,
                                                   double* __restrict__ work2)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   for (int m = 0; m < nc; m++) {
      long long i = m * a.ncg + c.c0;
      work2[i] = sig * kry_dir(x, i) + ycur[i];
   }
}

// out = x + (-gamma) (siginv jv + (-siginv) fcur); with pin.f also the FP32
// preconditioner right-hand side of out (the arithmetic of k_psolve_in), and
// out may then be null (the FP64 vector is not needed: the preconditioned
// vector replaces it)
__global__ void __launch_bounds__(256) k_gmres_finish(LevelArgs a, int nc, double siginv,
                                                      double gamma, KryDir x,
                                                      const double* __restrict__ jv,
                                                      const double* __restrict__ fcur,
                                                      double* __restrict__ out, PrecIn pin)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   for (int m = 0; m < nc; m++) {
      long long i = m * a.ncg + c.c0;
      double jvn = siginv * jv[i] + (-siginv) * fcur[i];
      double v = kry_dir(x, i) + (-gamma) * jvn;
      if (out) out[i] = v;
      if (pin.f) {
         bool q = m >= pin.iq && m < pin.iq + pin.ql;
         if (q)
            pin.f[pin.split ? m * a.ncg + split_at(a, c, 0) : i] = (float)(v / pin.sm[c.c0]);
         else
            pin.f[i] = (float)v;
      }
   }
}

// out = ewt work; partial = sum other * out (other may alias out)
__global__ void __launch_bounds__(256) k_scale_dot(LevelArgs a, int nc,
                                                   const double* __restrict__ ewt,
                                                   const double* __restrict__ work, double* out,
                                                   const double* other,
                                                   double* __restrict__ partial)
{
   CellPos c;
   double acc = 0.0;
   if (cell_pos(a, 0, c)) {
      const double w = red_w(a, c.c0);
      for (int m = 0; m < nc; m++) {
         long long i = m * a.ncg + c.c0;
         double v = ewt[i] * work[i];
         out[i] = v;
         double o = (other == out) ? v : other[i];
         acc += (o * v) * w;
      }
   }
   tile_store_sum(acc, partial);
}

// w = w + (-h) Vi; partial = sum Vnext * w (or w * w when Vnext is null)
__global__ void __launch_bounds__(256) k_mgs(LevelArgs a, int nc, double* __restrict__ w,
                                             const double* __restrict__ Vi, double h,
                                             const double* __restrict__ Vnext,
                                             double* __restrict__ partial)
{
   CellPos c;
   double acc = 0.0;
   if (cell_pos(a, 0, c)) {
      const double wc = red_w(a, c.c0);
      for (int m = 0; m < nc; m++) {
         long long i = m * a.ncg + c.c0;
         double v = w[i] + (-h) * Vi[i];
         w[i] = v;
         acc += (Vnext ? Vnext[i] * v : v * v) * wc;
      }
   }
   tile_store_sum(acc, partial);
}

// x = (sum_i y_i V_i) / ewt, summed in index order from zero
__global__ void __launch_bounds__(256) k_lincomb(LevelArgs a, int nc, VecList vl,
                                                 const double* __restrict__ ewt,
                                                 double* __restrict__ x)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   for (int m = 0; m < nc; m++) {
      long long i = m * a.ncg + c.c0;
      double acc = 0.0;
      for (int k = 0; k < vl.n; k++) acc = acc + vl.y[k] * vl.v[k][i];
      x[i] = acc / ewt[i];
   }
}

// b = gamma f - (y - a)
__global__ void __launch_bounds__(256) k_newton_rhs(LevelArgs a, int nc, double gamma,
                                                    const double* __restrict__ f,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ ah,
                                                    double* __restrict__ b)
{
   CellPos c;
   if (!cell_pos(a, 0, c)) return;
   for (int m = 0; m < nc; m++) {
      long long i = m * a.ncg + c.c0;
      b[i] = gamma * f[i] - (y[i] - ah[i]);
   }
}

// partial = sum (del ewt)^2; y = y + del
__global__ void __launch_bounds__(256) k_newton_update(LevelArgs a, int nc,
                                                       const double* __restrict__ del,
                                                       const double* __restrict__ ewt,
                                                       double* __restrict__ y,
                                                       double* __restrict__ partial)
{
   CellPos c;
   double acc = 0.0;
   if (cell_pos(a, 0, c)) {
      const double w = red_w(a, c.c0);
      for (int m = 0; m < nc; m++) {
         long long i = m * a.ncg + c.c0;
         double t = del[i] * ewt[i];
         acc += (t * t) * w;
         y[i] = y[i] + del[i];
      }
   }
   tile_store_sum(acc, partial);
}

// Per-tile sums of n values (n <= kMaxCgs) to partial[k * ntiles + tile].
__device__ __forceinline__ void tile_store_sums(const double* acc, int n, int ntiles,
                                                double* __restrict__ partial)
{
   __shared__ double sh[4][kMaxCgs];
   int t = threadIdx.x + TX * threadIdx.y;
#pragma unroll
   for (int k = 0; k < kMaxCgs; k++) {
      if (k >= n) break;
      double v = acc[k];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if ((t & 63) == 0) sh[t >> 6][k] = v;
   }
   __syncthreads();
   if (t < n) partial[(long long)t * ntiles + blockIdx.x] = (sh[0][t] + sh[1][t]) + (sh[2][t] + sh[3][t]);
}

__global__ void __launch_bounds__(256) k_cgs_dots(LevelArgs a, int nc, int ntiles,
                                                  const double* __restrict__ ewt,
                                                  const double* __restrict__ work, double* out,
                                                  PrecOut src, VecList vl,
                                                  double* __restrict__ partial)
{
   CellPos c;
   double acc[kMaxCgs];
#pragma unroll
   for (int k = 0; k < kMaxCgs; k++) acc[k] = 0.0;
   const int n = vl.n;
   if (cell_pos(a, 0, c)) {
      const double w = red_w(a, c.c0);
      for (int m = 0; m < nc; m++) {
         long long i = m * a.ncg + c.c0;
         double v;
         if (src.u) {
            // k_psolve_out's arithmetic, then the ewt scaling
            if (m >= src.iq && m < src.iq + src.ql) {
               v = (double)src.u[src.split ? m * a.ncg + split_at(a, c, 0) : i];
               v = v * src.sm[c.c0];
            } else {
               v = (double)src.u[i];
            }
            v = src.ewt[i] * v;
            out[i] = v;
         } else if (ewt) {
            v = ewt[i] * work[i];
            out[i] = v;
         } else {
            v = out[i];
         }
#pragma unroll
         for (int k = 0; k < kMaxCgs - 1; k++)
            if (k < n) acc[k] += (v * vl.v[k][i]) * w;
#pragma unroll
         for (int k = 0; k < kMaxCgs; k++)
            if (k == n) acc[k] += (v * v) * w;
      }
   }
   tile_store_sums(acc, n + 1, ntiles, partial);
}

__global__ void __launch_bounds__(256) k_cgs_update(LevelArgs a, int nc, double* __restrict__ w,
                                                    VecList vl, double* __restrict__ partial)
{
   CellPos c;
   double acc = 0.0;
   const int n = vl.n;
   if (cell_pos(a, 0, c)) {
      const double wc = red_w(a, c.c0);
      for (int m = 0; m < nc; m++) {
         long long i = m * a.ncg + c.c0;
         double v = w[i];
#pragma unroll
         for (int k = 0; k < kMaxCgs; k++)
            if (k < n) v = v + vl.y[k] * vl.v[k][i];
         w[i] = v;
         acc += (v * v) * wc;
      }
   }
   tile_store_sum(acc, partial);
}

#define AMPE_TILE_LAUNCH(KER, ...)                                                        \
   do {                                                                                   \
      hipLaunchKernelGGL(KER, dim3(L.ncell_tiles), dim3(TX, TY), 0, s, level_args(L),      \
                         __VA_ARGS__);                                                    \
      AMPE_HIP_CHECK(hipGetLastError());                                                  \
   } while (0)

int kry_prep(const Level& L, int nc, double hinv, double* V, const double* ewt, double* xk,
             double* sumsq, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_gmres_prep, nc, hinv, V, ewt, xk, L.partial);
   return reduce_partials(L, L.ncell_tiles, 1, 0, sumsq, s);
}

int kry_mid(const Level& L, int nc, double sig, const double* xk, const double* ycur,
            double* work2, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_gmres_mid, nc, sig, KryDir{xk, nullptr, 0.0, nullptr}, ycur, work2);
   return 0;
}

int kry_mid_scaled(const Level& L, int nc, double sig, const double* V, double h,
                   const double* ewt, const double* ycur, double* work2, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_gmres_mid, nc, sig, KryDir{nullptr, V, h, ewt}, ycur, work2);
   return 0;
}

int kry_finish(const Level& L, int nc, double siginv, double gamma, const double* xk,
               const double* jv, const double* fcur, double* out, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_gmres_finish, nc, siginv, gamma, KryDir{xk, nullptr, 0.0, nullptr}, jv,
                    fcur, out, PrecIn{});
   return 0;
}

int kry_finish_scaled(const Level& L, int nc, double siginv, double gamma, const double* V,
                      double h, const double* ewt, const double* jv, const double* fcur,
                      double* out, const PrecIn& pin, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_gmres_finish, nc, siginv, gamma, KryDir{nullptr, V, h, ewt}, jv, fcur,
                    out, pin);
   return 0;
}

int kry_scale_dot(const Level& L, int nc, const double* ewt, const double* work, double* out,
                  const double* other, double* dot, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_scale_dot, nc, ewt, work, out, other, L.partial);
   return reduce_partials(L, L.ncell_tiles, 1, 0, dot, s);
}

int kry_mgs(const Level& L, int nc, double* w, const double* Vi, double h, const double* Vnext,
            double* dot, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_mgs, nc, w, Vi, h, Vnext, L.partial);
   return reduce_partials(L, L.ncell_tiles, 1, 0, dot, s);
}

int kry_lincomb(const Level& L, int nc, const VecList& vl, const double* ewt, double* x,
                hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_lincomb, nc, vl, ewt, x);
   return 0;
}

int kry_newton_rhs(const Level& L, int nc, double gamma, const double* f, const double* y,
                   const double* ah, double* b, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_newton_rhs, nc, gamma, f, y, ah, b);
   return 0;
}

int kry_cgs_dots(const Level& L, int nc, const double* ewt, const double* work, double* out,
                 const VecList& vl, double* sums, hipStream_t s)
{
   if (vl.n + 1 > kMaxCgs) {
      set_error("kry_cgs_dots: too many vectors");
      return -1;
   }
   AMPE_TILE_LAUNCH(k_cgs_dots, nc, L.ncell_tiles, ewt, work, out, PrecOut{}, vl, L.partial);
   return reduce_regions(L, L.ncell_tiles, vl.n + 1, sums, s);
}

int kry_cgs_dots_from(const Level& L, int nc, const PrecOut& src, double* out,
                      const VecList& vl, double* sums, hipStream_t s)
{
   if (vl.n + 1 > kMaxCgs) {
      set_error("kry_cgs_dots: too many vectors");
      return -1;
   }
   AMPE_TILE_LAUNCH(k_cgs_dots, nc, L.ncell_tiles, nullptr, nullptr, out, src, vl, L.partial);
   return reduce_regions(L, L.ncell_tiles, vl.n + 1, sums, s);
}

int kry_cgs_update(const Level& L, int nc, double* w, const VecList& vl, double* sumsq,
                   hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_cgs_update, nc, w, vl, L.partial);
   return reduce_partials(L, L.ncell_tiles, 1, 0, sumsq, s);
}

int kry_newton_update(const Level& L, int nc, const double* del, const double* ewt, double* y,
                      double* sumsq, hipStream_t s)
{
   AMPE_TILE_LAUNCH(k_newton_update, nc, del, ewt, y, L.partial);
   return reduce_partials(L, L.ncell_tiles, 1, 0, sumsq, s);
}

}  // namespace ampe
