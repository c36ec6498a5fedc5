// Per-cell KKS phase-concentration solves.
//
// k_kks: the integrator's call over all ghosted cells of a level
// (PhaseConcentrationsStrategy::computePhaseConcentrations,
// source/PhaseConcentrationsStrategy.cc:26-127, ghost cells included per
// source/QuatIntegrator.cc:3119-3126), guess from the previous step's
// concentrations (source/CALPHADequilibriumPhaseConcentrationsStrategy.cc:
// 368-369) or from the surrogate MLP.
// k_kks_batch: independent cells with a per-cell temperature and the end
// members evaluated at that temperature, default start (0.5, 0.5)
// (tests/testCALPHADbinaryKKS.cc:123-127) or the surrogate.
//
// Each workgroup writes {sum of updates, max updates or failure marker,
// fallbacks, failures, surrogate starts} to its partial slot; a one-workgroup finish kernel folds them
// into device counters in a fixed order (no same-address atomics, no host
// synchronisation per call).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "kks.h"

namespace ampe {

constexpr double kKksFail = 1.0e9;
constexpr int kKksPart = 5;  // doubles per workgroup

// it: updates, mx: updates or kKksFail, fb: fallback (0/1), nf: failure
// (0/1), ss: surrogate start (0/1)
__device__ inline void kks_block_partial(double it, double mx, double fb, double nf, double ss,
                                         double* __restrict__ partial)
{
   for (int o = 32; o > 0; o >>= 1) {
      it += __shfl_xor(it, o);
      fb += __shfl_xor(fb, o);
      nf += __shfl_xor(nf, o);
      ss += __shfl_xor(ss, o);
      double v = __shfl_xor(mx, o);
      mx = v > mx ? v : mx;
   }
   __shared__ double sh[4 * kKksPart];
   int w = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 0) {
      sh[kKksPart * w] = it;
      sh[kKksPart * w + 1] = mx;
      sh[kKksPart * w + 2] = fb;
      sh[kKksPart * w + 3] = nf;
      sh[kKksPart * w + 4] = ss;
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      double s = 0.0, m = 0.0, f = 0.0, e = 0.0, g = 0.0;
      for (int k = 0; k < 4; k++) {
         s += sh[kKksPart * k];
         m = sh[kKksPart * k + 1] > m ? sh[kKksPart * k + 1] : m;
         f += sh[kKksPart * k + 2];
         e += sh[kKksPart * k + 3];
         g += sh[kKksPart * k + 4];
      }
      partial[kKksPart * blockIdx.x] = s;
      partial[kKksPart * blockIdx.x + 1] = m;
      partial[kKksPart * blockIdx.x + 2] = f;
      partial[kKksPart * blockIdx.x + 3] = e;
      partial[kKksPart * blockIdx.x + 4] = g;
   }
}

// Surrogate guess for the lane's cell; all 64 lanes of the wave call it.
__device__ inline void kks_surrogate_guess(const MlpDev& m, const u16x8* sf, const float* sb,
                                           double c, double phi, double T, double& gl,
                                           double& gs)
{
   float x[kMlpMaxIn] = {0.f, 0.f, 0.f, 0.f};
   x[0] = ((float)c - m.in_shift[0]) * m.in_scale[0];
   x[1] = ((float)phi - m.in_shift[1]) * m.in_scale[1];
   x[2] = ((float)T - m.in_shift[2]) * m.in_scale[2];
   float y[kMlpMaxOut];
   mlp_wave_infer<kMlpHidden>(sf, sb, m.nin, x, y);
   gl = (double)y[0] * m.out_scale[0] + m.out_shift[0];
   gs = (double)y[1] * m.out_scale[1] + m.out_shift[1];
}

// Every cell from the previous-step guess (kks_cell).
__global__ void __launch_bounds__(256) k_kks(long long n, ampe_model_params p,
                                             const double* __restrict__ phi,
                                             const double* __restrict__ c,
                                             const double* __restrict__ clr,
                                             const double* __restrict__ csr,
                                             double* __restrict__ cl, double* __restrict__ cs,
                                             double* __restrict__ partial)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   double it = 0.0, mx = 0.0;
   if (t < n) {
      double xl = clr[t], xs = csr[t];
      int r = kks_cell(p, c[t], phi[t], xl, xs);
      cl[t] = xl;
      cs[t] = xs;
      mx = r < 0 ? kKksFail : (double)r;
      it = r < 0 ? 0.0 : (double)r;
   }
   kks_block_partial(it, mx, 0.0, mx >= kKksFail ? 1.0 : 0.0, 0.0, partial);
}

// Switch policy, first pass (binary CALPHAD): a cell whose residual at the
// clamped previous-step guess is at most switch_tol (and switch_tol > 0) runs
// the Newton from that guess; the others are appended to the work list (one
// atomic per wave) and left to k_kks_surr.  Kept free of the network so that
// it runs at the occupancy of the plain solve.
__global__ void __launch_bounds__(256) k_kks_defer(long long n, ampe_model_params p,
                                                   double switch_tol,
                                                   const double* __restrict__ phi,
                                                   const double* __restrict__ c,
                                                   const double* __restrict__ clr,
                                                   const double* __restrict__ csr,
                                                   double* __restrict__ cl,
                                                   double* __restrict__ cs,
                                                   int* __restrict__ list,
                                                   unsigned* __restrict__ count,
                                                   double* __restrict__ partial)
{
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   double it = 0.0, mx = 0.0;
   bool bad = false;
   if (t < n) {
      const double T = p.temperature;
      double ct = c[t];
      double h = interp(phi[t], p.conc_interp);
      double xl = clr[t], xs = csr[t];
      kks_clamp(xl, xs);
      double f1, f2;
      kks_residual(p.calphad, T, 1.0 / (kRGas * T), h, ct, xl, xs, f1, f2);
      bad = switch_tol <= 0.0 || !(dmax(fabs(f1), fabs(f2)) <= switch_tol);
      if (!bad) {
         int r = kks_newton_from(p.calphad, T, h, ct, xl, xs, f1, f2, p.newton_max_its,
                                 p.newton_tol, -1, 0.0);
         cl[t] = xl;
         cs[t] = xs;
         mx = r < 0 ? kKksFail : (double)r;
         it = r < 0 ? 0.0 : (double)r;
      }
   }
   unsigned long long mask = __ballot(bad);
   if (mask) {
      const int lane = threadIdx.x & 63;
      const int first = __ffsll((long long)mask) - 1;
      unsigned base = 0;
      if (lane == first) base = atomicAdd(count, (unsigned)__popcll(mask));
      base = __shfl(base, first);
      if (bad) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int)t;
   }
   kks_block_partial(it, mx, 0.0, mx >= kKksFail ? 1.0 : 0.0, 0.0, partial);
}

// Switch policy, second pass: the listed cells start from the surrogate and
// restart from the previous-step guess when the residual after probe_its
// updates exceeds probe_tol.  Grid-stride over the list, whole waves at a
// time (the network's MFMA inference needs all 64 lanes).
__global__ void __launch_bounds__(256) k_kks_surr(ampe_model_params p, MlpDev m, int probe_its,
                                                  double probe_tol,
                                                  const double* __restrict__ phi,
                                                  const double* __restrict__ c,
                                                  const double* __restrict__ clr,
                                                  const double* __restrict__ csr,
                                                  double* __restrict__ cl,
                                                  double* __restrict__ cs,
                                                  const int* __restrict__ list,
                                                  const unsigned* __restrict__ count,
                                                  double* __restrict__ partial)
{
   using F = MlpFrags<kMlpHidden>;
   __shared__ u16x8 sf[F::NFWD * 64];
   __shared__ float sb[F::NBIAS];
   const long long nl = (long long)*count;
   double it = 0.0, mx = 0.0, fb = 0.0, ss = 0.0;
   if ((long long)blockIdx.x * blockDim.x < nl) {
      mlp_stage<kMlpHidden>(m, F::NFWD, sf, sb);
   }
   __syncthreads();
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const double T = p.temperature;
   for (long long base = ((long long)blockIdx.x * 4 + wave) * 64; base < nl;
        base += (long long)gridDim.x * blockDim.x) {
      long long i = base + lane;
      bool in = i < nl;
      long long t = in ? (long long)list[i] : 0;
      double ct = in ? c[t] : 0.5, pt = in ? phi[t] : 0.0;
      double gl, gs;
      kks_surrogate_guess(m, sf, sb, ct, pt, T, gl, gs);
      if (in) {
         double h = interp(pt, p.conc_interp);
         int r = kks_newton(p.calphad, T, h, ct, gl, gs, p.newton_max_its, p.newton_tol,
                            probe_its, probe_tol);
         if (r == -2) {
            fb += 1.0;
            double xl = clr[t], xs = csr[t];
            kks_clamp(xl, xs);
            double f1, f2;
            kks_residual(p.calphad, T, 1.0 / (kRGas * T), h, ct, xl, xs, f1, f2);
            int r2 = kks_newton_from(p.calphad, T, h, ct, xl, xs, f1, f2, p.newton_max_its,
                                     p.newton_tol, -1, 0.0);
            r = r2 < 0 ? -1 : r2 + probe_its;
            gl = xl;
            gs = xs;
         }
         cl[t] = gl;
         cs[t] = gs;
         ss += 1.0;
         if (r < 0) {
            mx = kKksFail;
         } else {
            it += (double)r;
            mx = (double)r > mx ? (double)r : mx;
         }
      }
   }
   kks_block_partial(it, mx, fb, mx >= kKksFail ? 1.0 : 0.0, ss, partial);
}

// acc[0] += updates, acc[1] = max(acc[1], max updates), acc[2] |= failure,
// acc[3] += 1 (calls), acc[4] += fallbacks, acc[5] += surrogate starts.
// One workgroup.
__global__ void __launch_bounds__(1024) k_kks_finish(const double* __restrict__ partial, int nb,
                                                     long long* __restrict__ acc,
                                                     unsigned* __restrict__ count)
{
   __shared__ double ss[1024], sm[1024], sf[1024], sg[1024];
   double s = 0.0, m = 0.0, f = 0.0, g = 0.0;
   for (int i = threadIdx.x; i < nb; i += 1024) {
      s += partial[kKksPart * i];
      m = partial[kKksPart * i + 1] > m ? partial[kKksPart * i + 1] : m;
      f += partial[kKksPart * i + 2];
      g += partial[kKksPart * i + 4];
   }
   ss[threadIdx.x] = s;
   sm[threadIdx.x] = m;
   sf[threadIdx.x] = f;
   sg[threadIdx.x] = g;
   __syncthreads();
   for (int w = 512; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
         ss[threadIdx.x] += ss[threadIdx.x + w];
         sm[threadIdx.x] = sm[threadIdx.x + w] > sm[threadIdx.x] ? sm[threadIdx.x + w]
                                                                 : sm[threadIdx.x];
         sf[threadIdx.x] += sf[threadIdx.x + w];
         sg[threadIdx.x] += sg[threadIdx.x + w];
      }
      __syncthreads();
   }
   if (threadIdx.x == 0) {
      bool fail = sm[0] >= kKksFail;
      acc[0] += (long long)ss[0];
      long long mxi = fail ? 0 : (long long)sm[0];
      acc[1] = mxi > acc[1] ? mxi : acc[1];
      if (fail) acc[2] = 1;
      acc[3] += 1;
      acc[4] += (long long)sf[0];
      acc[5] += (long long)sg[0];
      if (count) *count = 0u;  // the work list is empty again for the next call
   }
}

MlpDev Surrogate::dev() const
{
   MlpDev d;
   d.nin = nin;
   d.nout = nout;
   d.frag = frag;
   d.bias = bias;
   for (int i = 0; i < kMlpMaxIn; i++) {
      d.in_shift[i] = in_shift[i];
      d.in_scale[i] = in_scale[i];
   }
   for (int i = 0; i < kMlpMaxOut; i++) {
      d.out_shift[i] = out_shift[i];
      d.out_scale[i] = out_scale[i];
   }
   return d;
}

int launch_kks(const Level& L, const ampe_model_params& p, const double* phi,
               const double* c, const double* clr, const double* csr, double* cl,
               double* cs, const KksGuess& guess, long long* dev_acc, hipStream_t s)
{
   unsigned nb = (unsigned)((L.ncg + 255) / 256);
   unsigned nparts = nb;
   unsigned* count = nullptr;
   if (guess.surrogate && p.fe_model == AMPE_FE_CALPHAD_BINARY) {
      const Surrogate& S = *guess.surrogate;
      if (S.hid != kMlpHidden || S.nin != 3 || S.nout != 2) {
         set_error("launch_kks: the surrogate must map (c, phi, T) -> (cL, cS) with 64 hidden");
         return -1;
      }
      if (!guess.list || !guess.count) {
         set_error("launch_kks: the surrogate start needs a work list");
         return -1;
      }
      MlpDev m = S.dev();
      count = guess.count;
      hipLaunchKernelGGL(k_kks_defer, dim3(nb), dim3(256), 0, s, L.ncg, p, guess.switch_tol,
                         phi, c, clr, csr, cl, cs, guess.list, count, L.kks_part);
      AMPE_HIP_CHECK(hipGetLastError());
      unsigned g = std::min<unsigned>(kKksSurrGrid, nb);
      hipLaunchKernelGGL(k_kks_surr, dim3(g), dim3(256), 0, s, p, m, guess.probe_its,
                         guess.probe_tol, phi, c, clr, csr, cl, cs, guess.list, count,
                         L.kks_part + (size_t)kKksPart * nb);
      nparts = nb + g;
   } else {
      hipLaunchKernelGGL(k_kks, dim3(nb), dim3(256), 0, s, L.ncg, p, phi, c, clr, csr, cl, cs,
                         L.kks_part);
   }
   AMPE_HIP_CHECK(hipGetLastError());
   hipLaunchKernelGGL(k_kks_finish, dim3(1), dim3(1024), 0, s, L.kks_part, (int)nparts, dev_acc,
                      count);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int launch_kks_sync(const Level& L, const ampe_model_params& p, const double* phi,
                    const double* c, const double* clr, const double* csr, double* cl,
                    double* cs, int* max_iters, long long* total_iters, hipStream_t s)
{
   long long* acc = reinterpret_cast<long long*>(L.iscratch);
   AMPE_HIP_CHECK(hipMemsetAsync(acc, 0, sizeof(long long) * kKksAcc, s));
   if (launch_kks(L, p, phi, c, clr, csr, cl, cs, KksGuess{}, acc, s)) return -1;
   AMPE_HIP_CHECK(hipMemcpyAsync(L.hpinned, acc, sizeof(long long) * kKksAcc,
                                 hipMemcpyDeviceToHost, s));
   AMPE_HIP_CHECK(hipStreamSynchronize(s));
   long long h[kKksAcc];
   memcpy(h, L.hpinned, sizeof(h));
   if (max_iters) *max_iters = h[2] ? -1 : (int)h[1];
   if (total_iters) *total_iters = h[0];
   return 0;
}

// ------------------------------------------------------------ batch
// Grid of the batch kernel: at most this many workgroups, each walking its
// cells with a grid stride, so the per-workgroup partials stay bounded
// whatever the batch size (a 2^24-cell batch gives 8 cells per thread).
constexpr int kKksBatchMaxBlocks = 8192;

template <bool SURR>
__global__ void __launch_bounds__(256) k_kks_batch(long long n, ampe_model_params p,
                                                   ampe_calphad_tdep tab, MlpDev m,
                                                   int probe_its, double probe_tol,
                                                   const double* __restrict__ c,
                                                   const double* __restrict__ phi,
                                                   const double* __restrict__ T,
                                                   double* __restrict__ cl,
                                                   double* __restrict__ cs,
                                                   int* __restrict__ iters,
                                                   double* __restrict__ partial)
{
   using F = MlpFrags<kMlpHidden>;
   __shared__ u16x8 sf[SURR ? F::NFWD * 64 : 1];
   __shared__ float sb[SURR ? F::NBIAS : 1];
   if (SURR) {
      mlp_stage<kMlpHidden>(m, F::NFWD, sf, sb);
      __syncthreads();
   }
   // it: updates of converged cells, mx: largest count of a converged cell,
   // fb: fallbacks, nf: failures, ss: surrogate starts
   double it = 0.0, mx = 0.0, fb = 0.0, nf = 0.0, ss = 0.0;
   const long long stride = (long long)gridDim.x * blockDim.x;
   // the trip count is uniform over the workgroup, so every wave calls the
   // surrogate's wave-wide inference the same number of times
   for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += stride) {
      long long t = base + threadIdx.x;
      bool in = t < n;
      double ct = in ? c[t] : 0.5, pt = in ? phi[t] : 0.0, Tt = in ? T[t] : 1500.0;
      double gl = 0.5, gs = 0.5;
      if (SURR) kks_surrogate_guess(m, sf, sb, ct, pt, Tt, gl, gs);
      if (!in) continue;
      ampe_calphad_phase ph[2];
      calphad_at(tab, Tt, ph);
      double h = interp(pt, p.conc_interp);
      int r;
      if (SURR) {
         ss += 1.0;
         r = kks_newton(ph, Tt, h, ct, gl, gs, p.newton_max_its, p.newton_tol, probe_its,
                        probe_tol);
         if (r == -2) {
            fb += 1.0;
            gl = 0.5;
            gs = 0.5;
            int r2 = kks_newton(ph, Tt, h, ct, gl, gs, p.newton_max_its, p.newton_tol, -1, 0.0);
            r = r2 < 0 ? -1 : r2 + probe_its;
         }
      } else {
         r = kks_newton(ph, Tt, h, ct, gl, gs, p.newton_max_its, p.newton_tol, -1, 0.0);
      }
      cl[t] = gl;
      cs[t] = gs;
      if (iters) iters[t] = r;
      if (r < 0) {
         nf += 1.0;
      } else {
         it += (double)r;
         mx = (double)r > mx ? (double)r : mx;
      }
   }
   kks_block_partial(it, mx, fb, nf, ss, partial);
}

// Fixed-order fold of nb workgroup partials into {updates, largest count,
// fallbacks, failures}: each of 1024 lanes sums a fixed stride, then a fixed
// tree (the pattern of k_kks_finish).
__global__ void __launch_bounds__(1024) k_kks_batch_finish(const double* __restrict__ partial,
                                                           int nb, long long* __restrict__ acc)
{
   __shared__ double ss[1024], sm[1024], sf[1024], se[1024];
   double s = 0.0, m = 0.0, f = 0.0, e = 0.0;
   for (int i = threadIdx.x; i < nb; i += 1024) {
      s += partial[kKksPart * i];
      m = partial[kKksPart * i + 1] > m ? partial[kKksPart * i + 1] : m;
      f += partial[kKksPart * i + 2];
      e += partial[kKksPart * i + 3];
   }
   ss[threadIdx.x] = s;
   sm[threadIdx.x] = m;
   sf[threadIdx.x] = f;
   se[threadIdx.x] = e;
   __syncthreads();
   for (int w = 512; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
         ss[threadIdx.x] += ss[threadIdx.x + w];
         sm[threadIdx.x] = sm[threadIdx.x + w] > sm[threadIdx.x] ? sm[threadIdx.x + w]
                                                                 : sm[threadIdx.x];
         sf[threadIdx.x] += sf[threadIdx.x + w];
         se[threadIdx.x] += se[threadIdx.x + w];
      }
      __syncthreads();
   }
   if (threadIdx.x == 0) {
      acc[0] = (long long)ss[0];
      acc[1] = (long long)sm[0];
      acc[2] = (long long)sf[0];
      acc[3] = (long long)se[0];
   }
}

int kks_batch_check(const ampe_model_params& p, const ampe_calphad_tdep& tab, long long n,
                    const Surrogate* guess)
{
   if (n < 0) {
      set_error("kks_batch: n must be >= 0");
      return -1;
   }
   if (p.fe_model != AMPE_FE_CALPHAD_BINARY) {
      set_error("kks_batch: the model must be binary CALPHAD");
      return -1;
   }
   for (int ph = 0; ph < 2; ph++)
      for (int sp = 0; sp < 2; sp++) {
         int nr = tab.g[ph][sp].nrange;
         if (nr < 1 || nr > 4) {
            set_error("kks_batch: gfun nrange must be in [1, 4]");
            return -1;
         }
      }
   if (guess && (guess->hid != kMlpHidden || guess->nin != 3 || guess->nout != 2)) {
      set_error("kks_batch: the surrogate must map (c, phi, T) -> (cL, cS) with 64 hidden");
      return -1;
   }
   return 0;
}

int kks_batch_blocks(long long n)
{
   long long nbl = (n + 255) / 256;
   return (int)std::min<long long>(nbl, kKksBatchMaxBlocks);
}

int kks_batch_launch(const ampe_model_params& p, const ampe_calphad_tdep& tab, long long n,
                     const double* c, const double* phi, const double* T, const Surrogate* guess,
                     int probe_its, double probe_tol, double* cl, double* cs, int* iters,
                     double* part, long long* acc, hipStream_t s)
{
   const int nb = kks_batch_blocks(n);
   MlpDev m{};
   if (guess) {
      m = guess->dev();
      hipLaunchKernelGGL(k_kks_batch<true>, dim3((unsigned)nb), dim3(256), 0, s, n, p, tab, m,
                         probe_its, probe_tol, c, phi, T, cl, cs, iters, part);
   } else {
      hipLaunchKernelGGL(k_kks_batch<false>, dim3((unsigned)nb), dim3(256), 0, s, n, p, tab, m,
                         -1, 0.0, c, phi, T, cl, cs, iters, part);
   }
   AMPE_HIP_CHECK(hipGetLastError());
   hipLaunchKernelGGL(k_kks_batch_finish, dim3(1), dim3(1024), 0, s, part, nb, acc);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

static void kks_batch_stats_of(const long long h[4], ampe_kks_batch_stats* stats)
{
   if (!stats) return;
   stats->iters = h[0];
   stats->max_iters = h[1];
   stats->fallbacks = h[2];
   stats->failures = h[3];
}

int kks_batch(const ampe_model_params& p, const ampe_calphad_tdep& tab, long long n,
              const double* c, const double* phi, const double* T, const Surrogate* guess,
              int probe_its, double probe_tol, double* cl, double* cs, int* iters,
              ampe_kks_batch_stats* stats, hipStream_t s)
{
   if (kks_batch_check(p, tab, n, guess)) return -1;
   if (n == 0) {
      if (stats) std::memset(stats, 0, sizeof(*stats));
      return 0;
   }
   const int nb = kks_batch_blocks(n);
   double* part = nullptr;
   long long* acc = nullptr;
   AMPE_HIP_CHECK(hipMalloc((void**)&part, sizeof(double) * kKksPart * nb));
   if (hipMalloc((void**)&acc, sizeof(long long) * 4) != hipSuccess) {
      (void)hipFree(part);
      set_error("kks_batch: out of device memory");
      return -1;
   }
   int rc = kks_batch_launch(p, tab, n, c, phi, T, guess, probe_its, probe_tol, cl, cs, iters,
                             part, acc, s);
   long long h[4] = {0, 0, 0, 0};
   hipError_t e = hipSuccess;
   if (rc == 0) e = hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s);
   if (rc == 0 && e == hipSuccess) e = hipStreamSynchronize(s);
   (void)hipFree(part);
   (void)hipFree(acc);
   if (rc) return rc;
   if (e != hipSuccess) {
      set_error(std::string("kks_batch: ") + hipGetErrorString(e));
      return -1;
   }
   kks_batch_stats_of(h, stats);
   return 0;
}

// ------------------------------------------------------------ batch without allocation
// The workgroup partials of the batch live in one file-scope device array per
// GPU: kKksPart doubles for each of at most kKksBatchMaxBlocks workgroups (the
// batch grid never has more, whatever n), then four 8-byte counters that
// stand in for the caller's statistics when it passes none.  In host code the
// array's name is not a device address; the address comes from
// hipGetSymbolAddress, once per device with that device current.
constexpr long long kKksBatchScratchDoubles = (long long)kKksPart * kKksBatchMaxBlocks + 4;
static __device__ __align__(256) double scratch[kKksBatchScratchDoubles];

constexpr int kKksScratchMaxDevices = 64;
static std::mutex g_kks_scratch_mu;
static double* g_kks_scratch[kKksScratchMaxDevices] = {};

static int kks_batch_scratch(double** out)
{
   int dev = -1;
   AMPE_HIP_CHECK(hipGetDevice(&dev));
   if (dev < 0 || dev >= kKksScratchMaxDevices) {
      set_error("kks_batch_noalloc: device ordinal out of range");
      return -1;
   }
   std::lock_guard<std::mutex> lk(g_kks_scratch_mu);
   if (!g_kks_scratch[dev]) {
      void* ptr = nullptr;
      hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr) {
         set_error(std::string("kks_batch_noalloc: hipGetSymbolAddress: ") +
                   hipGetErrorString(e));
         return -1;
      }
      g_kks_scratch[dev] = static_cast<double*>(ptr);
   }
   *out = g_kks_scratch[dev];
   return 0;
}

int kks_batch_noalloc(const ampe_model_params& p, const ampe_calphad_tdep& tab, long long n,
                      const double* c, const double* phi, const double* T,
                      const Surrogate* guess, int probe_its, double probe_tol, double* cl,
                      double* cs, int* iters, ampe_kks_batch_stats* stats_dev, hipStream_t s)
{
   if (kks_batch_check(p, tab, n, guess)) return -1;
   if (kks_batch_blocks(n) > kKksBatchMaxBlocks) {
      set_error("kks_batch_noalloc: grid exceeds the scratch partials");
      return -1;
   }
   double* base = nullptr;
   if (kks_batch_scratch(&base)) return -1;
   long long* acc = stats_dev ? reinterpret_cast<long long*>(stats_dev)
                              : reinterpret_cast<long long*>(base + kKksPart * kKksBatchMaxBlocks);
   if (n == 0) {
      // zero counters: the finish kernel over no partials writes zeros
      hipLaunchKernelGGL(k_kks_batch_finish, dim3(1), dim3(1024), 0, s, base, 0, acc);
      AMPE_HIP_CHECK(hipGetLastError());
      return 0;
   }
   return kks_batch_launch(p, tab, n, c, phi, T, guess, probe_its, probe_tol, cl, cs, iters,
                           base, acc, s);
}

}  // namespace ampe
