// Vector kernels of the Newton-Krylov loop: element-wise updates over whole
// (ghosted) vectors and deterministic interior reductions for the SPGMR dot
// products and CVODE WRMS norms (Sundials_SAMRAIVector::dotWith /
// weightedRMSNorm, source/samrai/Sundials_SAMRAIVector.cc:156-177).
#include "fac.h"

namespace ampe {

static unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

#define AMPE_FLAT_KERNEL(NAME, PARAMS, BODY)                                      \
   __global__ void __launch_bounds__(256) NAME PARAMS                             \
   {                                                                              \
      long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;             \
      if (t >= n) return;                                                         \
      BODY;                                                                       \
   }

AMPE_FLAT_KERNEL(k_axpby, (long long n, double a, const double* __restrict__ x, double b,
                           const double* __restrict__ y, double* __restrict__ z),
                 z[t] = a * x[t] + b * y[t])
AMPE_FLAT_KERNEL(k_fill, (long long n, double v, double* __restrict__ z), z[t] = v)
AMPE_FLAT_KERNEL(k_mul, (long long n, const double* __restrict__ x, const double* __restrict__ y,
                         double* __restrict__ z),
                 z[t] = x[t] * y[t])
AMPE_FLAT_KERNEL(k_div, (long long n, const double* __restrict__ x, const double* __restrict__ y,
                         double* __restrict__ z),
                 z[t] = x[t] / y[t])
AMPE_FLAT_KERNEL(k_ewt, (long long n, const double* __restrict__ y, double rtol, double atol,
                         double* __restrict__ w),
                 w[t] = 1.0 / (rtol * fabs(y[t]) + atol))
AMPE_FLAT_KERNEL(k_hist, (long long n, double c1, const double* __restrict__ y, double c2,
                          const double* __restrict__ yp, double* __restrict__ o),
                 o[t] = c1 * y[t] - c2 * yp[t])
AMPE_FLAT_KERNEL(k_to_f32, (long long n, const double* __restrict__ x, float* __restrict__ y),
                 y[t] = (float)x[t])
AMPE_FLAT_KERNEL(k_to_f64, (long long n, const float* __restrict__ x, double* __restrict__ y),
                 y[t] = (double)x[t])
AMPE_FLAT_KERNEL(k_add_f, (long long n, float* __restrict__ u, const float* __restrict__ e),
                 u[t] = u[t] + e[t])

template <typename T>
__global__ void __launch_bounds__(256) k_zero(long long n, T* __restrict__ z)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t < n) z[t] = T(0);
}

// A kernel, not hipMemsetAsync: the preconditioner solve may be captured
// into a HIP graph, and its buffers are cleared by kernel nodes there.
template <typename T>
int fill_zero(T* z, long long n, hipStream_t s)
{
   if (n <= 0) return 0;
   hipLaunchKernelGGL(k_zero<T>, dim3(nblk(n)), dim3(256), 0, s, n, z);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}
template int fill_zero<float>(float*, long long, hipStream_t);
template int fill_zero<double>(double*, long long, hipStream_t);
AMPE_FLAT_KERNEL(k_sqrt, (long long n, const double* __restrict__ x, double* __restrict__ y),
                 y[t] = sqrt(x[t]))

#define AMPE_FLAT_LAUNCH(KER, ...)                                                \
   do {                                                                           \
      if (n <= 0) return 0;                                                       \
      hipLaunchKernelGGL(KER, dim3(nblk(n)), dim3(256), 0, s, n, __VA_ARGS__);    \
      AMPE_HIP_CHECK(hipGetLastError());                                          \
      return 0;                                                                   \
   } while (0)

int vec_axpby(long long n, double a, const double* x, double b, const double* y, double* z,
              hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_axpby, a, x, b, y, z);
}
int vec_fill(long long n, double v, double* z, hipStream_t s) { AMPE_FLAT_LAUNCH(k_fill, v, z); }
int vec_mul(long long n, const double* x, const double* y, double* z, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_mul, x, y, z);
}
int vec_div(long long n, const double* x, const double* y, double* z, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_div, x, y, z);
}
int vec_ewt(long long n, const double* y, double rtol, double atol, double* w, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_ewt, y, rtol, atol, w);
}
int vec_bdf2_hist(long long n, double c1, const double* y, double c2, const double* yp,
                  double* out, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_hist, c1, y, c2, yp, out);
}
int vec_to_f32(long long n, const double* x, float* y, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_to_f32, x, y);
}
int vec_axpby_f(long long n, float* u, const float* e, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_add_f, u, e);
}
int vec_to_f64(long long n, const float* x, double* y, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_to_f64, x, y);
}
int vec_div_f(long long n, const double* x, const double* sm, double* y, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_div, x, sm, y);
}
int vec_sqrt(long long n, const double* x, double* y, hipStream_t s)
{
   AMPE_FLAT_LAUNCH(k_sqrt, x, y);
}

// Per-tile partial of sum_m sum_cells x*y (mode 0) or (x*w)^2 (mode 1) over
// interior cells; one partial per tile, reduced in fixed order afterwards.
__global__ void __launch_bounds__(256) k_dot_partial(LevelArgs a, int ncomp, int mode,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     double* __restrict__ partial)
{
   CellPos c;
   double acc = 0.0;
   if (cell_pos(a, 0, c)) {
      const double w = red_w(a, c.c0);
      for (int m = 0; m < ncomp; m++) {
         long long i = m * a.ncg + c.c0;
         double v = (mode == 0) ? x[i] * y[i] : (x[i] * y[i]) * (x[i] * y[i]);
         acc += v * w;
      }
   }
   // wave reduction then across the 4 waves of the tile
   for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
   __shared__ double sh[4];
   int t = threadIdx.x + TX * threadIdx.y;
   if ((t & 63) == 0) sh[t >> 6] = acc;
   __syncthreads();
   if (t == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

int vec_dot(const Level& L, int ncomp, const double* x, const double* y, double* out,
            hipStream_t s)
{
   hipLaunchKernelGGL(k_dot_partial, dim3(L.ncell_tiles), dim3(TX, TY), 0, s, level_args(L),
                      ncomp, 0, x, y, L.partial);
   AMPE_HIP_CHECK(hipGetLastError());
   return reduce_partials(L, L.ncell_tiles, 1, 0, out, s);
}

int vec_sumsq_w(const Level& L, int ncomp, const double* x, const double* w, double* out,
                hipStream_t s)
{
   hipLaunchKernelGGL(k_dot_partial, dim3(L.ncell_tiles), dim3(TX, TY), 0, s, level_args(L),
                      ncomp, 1, x, w, L.partial);
   AMPE_HIP_CHECK(hipGetLastError());
   return reduce_partials(L, L.ncell_tiles, 1, 0, out, s);
}

// Preconditioner input over all nc ghosted component blocks: u = 0 and
// f = r, with the orientation blocks [iq, iq + ql) divided by sqrt(M) (FP64
// division, then the conversion to T).  For T = double f may be null: then
// only the orientation blocks are written, to fq (ql blocks).  np > 0: the
// orientation blocks are stored in the row parity-split layout of the level
// whose patch table is P (np patches).
template <typename T>
__global__ void __launch_bounds__(256) k_psolve_in(long long ng, int nc, int iq, int ql,
                                                   const double* __restrict__ r,
                                                   const double* __restrict__ sm,
                                                   T* __restrict__ f, T* __restrict__ fq,
                                                   T* __restrict__ u, const PatchDev* P, int np)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= ng * nc) return;
   long long m = i / ng, c = i - m * ng;
   u[i] = T(0);
   bool q = m >= iq && m < iq + ql;
   if (q) {
      double v = r[i] / sm[c];
      long long o = np > 0 ? m * ng + split_index(P, np, c) : i;
      if (f) f[o] = (T)v;
      else fq[o - (long long)iq * ng] = (T)v;
   } else if (f) {
      f[i] = (T)r[i];
   }
}

// Preconditioner output: z = u, orientation blocks multiplied by sqrt(M),
// then by ewt when given (z = ewt u).
template <typename T>
__global__ void __launch_bounds__(256) k_psolve_out(long long ng, int nc, int iq, int ql,
                                                    const T* __restrict__ u,
                                                    const double* __restrict__ sm,
                                                    const double* __restrict__ ewt,
                                                    double* __restrict__ z, const PatchDev* P,
                                                    int np)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= ng * nc) return;
   long long m = i / ng, c = i - m * ng;
   double v;
   if (m >= iq && m < iq + ql) {
      v = (double)u[np > 0 ? m * ng + split_index(P, np, c) : i];
      v = v * sm[c];
   } else {
      v = (double)u[i];
   }
   if (ewt) v = ewt[i] * v;
   z[i] = v;
}

template <typename T>
int psolve_in(long long ng, int nc, int iq, int ql, const double* r, const double* sm, T* f,
              T* fq, T* u, hipStream_t s, const Level* split)
{
   long long n = ng * nc;
   hipLaunchKernelGGL(k_psolve_in<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ng, nc,
                      iq, ql, r, sm, f, fq, u, split ? split->dp : nullptr,
                      split ? split->npatch : 0);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template <typename T>
int psolve_out(long long ng, int nc, int iq, int ql, const T* u, const double* sm,
               const double* ewt, double* z, hipStream_t s, const Level* split)
{
   long long n = ng * nc;
   hipLaunchKernelGGL(k_psolve_out<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ng,
                      nc, iq, ql, u, sm, ewt, z, split ? split->dp : nullptr,
                      split ? split->npatch : 0);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

// x[m][c] *= sm[c] for ncomp blocks of ng entries
__global__ void __launch_bounds__(256) k_scale_comp(long long ng, int ncomp,
                                                    const double* __restrict__ sm,
                                                    double* __restrict__ x)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= ng * ncomp) return;
   long long m = i / ng;
   x[i] = x[i] * sm[i - m * ng];
}

int vec_scale_comp(long long ng, int ncomp, const double* sm, double* x, hipStream_t s)
{
   long long n = ng * ncomp;
   if (n <= 0) return 0;
   hipLaunchKernelGGL(k_scale_comp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ng, ncomp,
                      sm, x);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

template int psolve_in<float>(long long, int, int, int, const double*, const double*, float*,
                              float*, float*, hipStream_t, const Level*);
template int psolve_in<double>(long long, int, int, int, const double*, const double*, double*,
                               double*, double*, hipStream_t, const Level*);
template int psolve_out<float>(long long, int, int, int, const float*, const double*,
                               const double*, double*, hipStream_t, const Level*);
template int psolve_out<double>(long long, int, int, int, const double*, const double*,
                                const double*, double*, hipStream_t, const Level*);

}  // namespace ampe
