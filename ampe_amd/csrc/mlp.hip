// KKS surrogate MLP: parameter packing into MFMA fragments, batched
// inference, loss gradient (forward + backward on the matrix cores), Adam and
// the training loop.  Layout conventions are described in mlp.h.
#include <cmath>
#include <cstring>
#include <vector>

#include "level.h"
#include "mlp.h"

namespace ampe {

// Element e of the fragment image: fragment f = e / 512, lane = (e % 512) / 8,
// K slot i = e % 8 of that lane; the value is the BF16 of the FP32 master
// parameter the slot stands for (zero for padding).  Threads beyond the image
// write the FP32 bias array (b1, b2, b3 padded to 16).
template <int H>
__global__ void __launch_bounds__(256) k_mlp_pack(const float* __restrict__ w, int nin, int nout,
                                                  unsigned short* __restrict__ frag,
                                                  float* __restrict__ bias)
{
   using L = MlpFrags<H>;
   const long long ob1 = (long long)H * nin, ow2 = ob1 + H, ob2 = ow2 + (long long)H * H;
   const long long ow3 = ob2 + H, ob3 = ow3 + (long long)nout * H;
   int e = blockIdx.x * blockDim.x + threadIdx.x;
   if (e < L::NALL * 512) {
      int f = e >> 9, lane = (e >> 3) & 63, i = e & 7;
      int g = lane >> 4, r = lane & 15, s = 8 * g + i;
      float v = 0.0f;
      if (f < L::F2) {  // W1: row = hidden unit, slot = input feature
         int hu = 16 * (f - L::F1) + r;
         if (s < nin) v = w[(long long)hu * nin + s];
      } else if (f < L::F3) {  // W2: row = output unit, slot -> input unit
         int q = f - L::F2, mt = q / L::KB, kb = q % L::KB;
         v = w[ow2 + (long long)(16 * mt + r) * H + 32 * kb + mlp_perm(s)];
      } else if (f < L::G2) {  // W3: row = output, slot -> hidden unit
         int kb = f - L::F3;
         if (r < nout) v = w[ow3 + (long long)r * H + 32 * kb + mlp_perm(s)];
      } else if (f < L::G3) {  // W2^T: row = input unit, slot -> output unit
         int q = f - L::G2, mt = q / L::KB, kb = q % L::KB;
         v = w[ow2 + (long long)(32 * kb + mlp_perm(s)) * H + 16 * mt + r];
      } else {  // W3^T: row = hidden unit, slot -> output
         int mt = f - L::G3, o = mlp_perm(s);
         if (o < nout) v = w[ow3 + (long long)o * H + 16 * mt + r];
      }
      frag[e] = f2bf(v);
   } else if (e < L::NALL * 512 + L::NBIAS) {
      int b = e - L::NALL * 512;
      float v;
      if (b < H) v = w[ob1 + b];
      else if (b < 2 * H) v = w[ob2 + b - H];
      else v = (b - 2 * H < nout) ? w[ob3 + b - 2 * H] : 0.0f;
      bias[b] = v;
   }
}

// Inference on normalised inputs x[nin][n] -> y[nout][n].
template <int H>
__global__ void __launch_bounds__(256) k_mlp_forward(MlpDev m, long long n,
                                                     const float* __restrict__ x,
                                                     float* __restrict__ y)
{
   using L = MlpFrags<H>;
   __shared__ u16x8 sf[L::NFWD * 64];
   __shared__ float sb[L::NBIAS];
   mlp_stage<H>(m, L::NFWD, sf, sb);
   __syncthreads();
   long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   float xi[kMlpMaxIn] = {0.f, 0.f, 0.f, 0.f};
   for (int i = 0; i < m.nin; i++) xi[i] = t < n ? x[i * n + t] : 0.0f;
   float yo[kMlpMaxOut];
   mlp_wave_infer<H>(sf, sb, m.nin, xi, yo);
   if (t < n)
      for (int o = 0; o < m.nout; o++) y[o * n + t] = yo[o];
}

__device__ inline f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c)
{
   return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Write a result-layout tile (lane (g, r) holds rows 4g..4g+3 of column r)
// to the wave's LDS scratch and read it back transposed as BF16: lane (g, r)
// receives T[r][4g..4g+3], the A or B operand of a 16x16x16 product whose K
// index is the tile's column (the sample).  Called by every thread of the
// workgroup the same number of times.
__device__ inline s16x4 tile_T(float* sc, f32x4 v)
{
   const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
   __syncthreads();
   for (int i = 0; i < 4; i++) sc[(4 * g + i) * 17 + r] = v[i];
   __syncthreads();
   s16x4 o;
   for (int i = 0; i < 4; i++) o[i] = (short)f2bf(sc[r * 17 + 4 * g + i]);
   return o;
}

__device__ inline float sum16(float v)
{
   v += __shfl_xor(v, 1);
   v += __shfl_xor(v, 2);
   v += __shfl_xor(v, 4);
   v += __shfl_xor(v, 8);
   return v;
}

// Loss and gradient over samples [base, base + n) of x[nin][ld], t[nout][ld].
// Each wave takes tpw consecutive 16-sample tiles and writes one partial row
// of ncol = nparam + 1 floats (gradient, then its sum of squared errors).
template <int H>
__global__ void __launch_bounds__(256) k_mlp_grad(MlpDev m, long long n, long long base,
                                                  long long ld, int tpw,
                                                  const float* __restrict__ x,
                                                  const float* __restrict__ t, float inv_n,
                                                  float* __restrict__ partial, int ncol)
{
   using L = MlpFrags<H>;
   constexpr int MT = L::MT, KB = L::KB;
   extern __shared__ __align__(16) unsigned char smem[];
   u16x8* sf = reinterpret_cast<u16x8*>(smem);
   float* sb = reinterpret_cast<float*>(sf + L::NALL * 64);
   float* sc = sb + L::NBIAS + (threadIdx.x >> 6) * (16 * 17);
   mlp_stage<H>(m, L::NALL, sf, sb);
   __syncthreads();
   const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
   const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
   const int nin = m.nin, nout = m.nout;
   const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
   f32x4 a1[MT], a2[MT][MT], a3[MT], db1[MT], db2[MT], db3 = zero;
   for (int i = 0; i < MT; i++) {
      a1[i] = a3[i] = db1[i] = db2[i] = zero;
      for (int j = 0; j < MT; j++) a2[i][j] = zero;
   }
   float lsum = 0.0f;
   for (int k = 0; k < tpw; k++) {
      long long s = (wave * tpw + k) * 16 + r;
      bool valid = s < n;
      f32x4 xt = zero;
      if (g == 0 && valid)
         for (int i = 0; i < nin; i++) xt[i] = x[i * ld + base + s];
      u16x8 xu;
      for (int i = 0; i < 4; i++) {
         xu[i] = f2bf(xt[i]);
         xu[i + 4] = 0;
      }
      f32x4 h1[MT], h2[MT], y;
      mlp_tile_forward<H>(sf, sb, as_bf(xu), h1, h2, y);
      // dL/dy = 2 (y - t) / n
      f32x4 dy = zero;
      if (g == 0 && valid)
         for (int o = 0; o < nout; o++) {
            float e = y[o] - t[o * ld + base + s];
            lsum += e * e;
            dy[o] = 2.0f * e * inv_n;
         }
      db3 += dy;
      // dW3 += dY^T H2
      s16x4 tdy = tile_T(sc, dy);
#pragma unroll
      for (int mt = 0; mt < MT; mt++) a3[mt] = mfma16(tdy, tile_T(sc, h2[mt]), a3[mt]);
      // dZ2 = (W3^T dY^T) * (1 - h2^2)
      bf16x8 dyb = pack8(dy, zero);
      f32x4 dz2[MT];
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
         f32x4 acc = mfma32(sf[(L::G3 + mt) * 64 + lane], dyb, zero);
         for (int i = 0; i < 4; i++) dz2[mt][i] = acc[i] * (1.0f - h2[mt][i] * h2[mt][i]);
         db2[mt] += dz2[mt];
      }
      // dW2 += dZ2^T H1
      s16x4 tz[MT], th[MT];
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
         tz[mt] = tile_T(sc, dz2[mt]);
         th[mt] = tile_T(sc, h1[mt]);
      }
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
         for (int kt = 0; kt < MT; kt++) a2[mt][kt] = mfma16(tz[mt], th[kt], a2[mt][kt]);
      // dZ1 = (W2^T dZ2^T) * (1 - h1^2)
      bf16x8 zb[KB];
#pragma unroll
      for (int kb = 0; kb < KB; kb++) zb[kb] = pack8(dz2[2 * kb], dz2[2 * kb + 1]);
      s16x4 tx = tile_T(sc, xt);
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
         f32x4 acc = zero;
#pragma unroll
         for (int kb = 0; kb < KB; kb++)
            acc = mfma32(sf[(L::G2 + mt * KB + kb) * 64 + lane], zb[kb], acc);
         f32x4 dz1;
         for (int i = 0; i < 4; i++) dz1[i] = acc[i] * (1.0f - h1[mt][i] * h1[mt][i]);
         db1[mt] += dz1;
         // dW1 += dZ1^T X
         a1[mt] = mfma16(tile_T(sc, dz1), tx, a1[mt]);
      }
   }
   // partial row of this wave
   const long long ob1 = (long long)H * nin, ow2 = ob1 + H, ob2 = ow2 + (long long)H * H;
   const long long ow3 = ob2 + H, ob3 = ow3 + (long long)nout * H;
   float* row = partial + wave * ncol;
#pragma unroll
   for (int mt = 0; mt < MT; mt++)
      for (int i = 0; i < 4; i++) {
         int u = 16 * mt + 4 * g + i;
         if (r < nin) row[(long long)u * nin + r] = a1[mt][i];
#pragma unroll
         for (int kt = 0; kt < MT; kt++) row[ow2 + (long long)u * H + 16 * kt + r] = a2[mt][kt][i];
         int o = 4 * g + i;
         if (o < nout) row[ow3 + (long long)o * H + 16 * mt + r] = a3[mt][i];
         float b1 = sum16(db1[mt][i]), b2 = sum16(db2[mt][i]);
         if (r == 0) {
            row[ob1 + u] = b1;
            row[ob2 + u] = b2;
         }
      }
   for (int i = 0; i < 4; i++) {
      float b3 = sum16(db3[i]);
      if (r == 0 && 4 * g + i < nout) row[ob3 + 4 * g + i] = b3;
   }
   for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
   if (lane == 0) row[ncol - 1] = lsum;
}

// Column sums of nrows partial rows (FP64): grad[j] for j < ncol - 1; the
// last column gives the loss (times inv_n).  A workgroup owns 64 columns
// (one per lane, so every row read is one coalesced line) and its kRedWaves
// waves split the rows round robin, eight independent loads in flight per
// lane; the waves' sums are then added in wave order, so the result does not
// depend on timing.
constexpr int kRedWaves = 16;

__global__ void __launch_bounds__(64 * kRedWaves) k_mlp_reduce(const float* __restrict__ partial,
                                                              int nrows, int ncol, double inv_n,
                                                              float* __restrict__ grad,
                                                              double* __restrict__ loss)
{
   __shared__ double sh[kRedWaves][64];
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   const int j = blockIdx.x * 64 + lane;
   double s = 0.0;
   if (j < ncol) {
      int i = w;
      for (; i + 7 * kRedWaves < nrows; i += 8 * kRedWaves) {
         float v[8];
#pragma unroll
         for (int u = 0; u < 8; u++) v[u] = partial[(long long)(i + u * kRedWaves) * ncol + j];
#pragma unroll
         for (int u = 0; u < 8; u++) s += (double)v[u];
      }
      for (; i < nrows; i += kRedWaves) s += (double)partial[(long long)i * ncol + j];
   }
   sh[w][lane] = s;
   __syncthreads();
   if (w == 0 && j < ncol) {
      double t = 0.0;
      for (int k = 0; k < kRedWaves; k++) t += sh[k][lane];
      if (j < ncol - 1)
         grad[j] = (float)t;
      else
         *loss = t * inv_n;
   }
}

// Adam (Kingma & Ba 2015) with bias corrections bc1 = 1 - beta1^t,
// bc2 = 1 - beta2^t.
__global__ void __launch_bounds__(256) k_mlp_adam(long long np, float* __restrict__ w,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ g, float lr,
                                                  float b1, float b2, float eps, float bc1,
                                                  float bc2)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= np) return;
   float gi = g[i];
   float mi = b1 * m[i] + (1.0f - b1) * gi;
   float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
   m[i] = mi;
   v[i] = vi;
   float mh = mi / bc1, vh = vi / bc2;
   w[i] = w[i] - lr * mh / (sqrtf(vh) + eps);
}

// out[f][i] = float(raw[f][i] - shift_f) * scale_f  (inputs: FP32 after the
// cast, as in the KKS kernels) or float((raw - shift) / scale) (targets).
__global__ void __launch_bounds__(256) k_mlp_norm(long long n, int nf, int target,
                                                  const double* __restrict__ raw, MlpDev m,
                                                  float* __restrict__ out)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   for (int f = 0; f < nf; f++) {
      double v = raw[f * n + i];
      out[f * n + i] = target ? (float)((v - m.out_shift[f]) / m.out_scale[f])
                              : ((float)v - m.in_shift[f]) * m.in_scale[f];
   }
}

// ------------------------------------------------------------ host
static int pack(Surrogate* S, hipStream_t st)
{
   using L = MlpFrags<kMlpHidden>;
   int total = L::NALL * 512 + L::NBIAS;
   hipLaunchKernelGGL(k_mlp_pack<kMlpHidden>, dim3((total + 255) / 256), dim3(256), 0, st, S->w,
                      S->nin, S->nout, reinterpret_cast<unsigned short*>(S->frag), S->bias);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

// splitmix64 (Steele, Lea & Flood 2014)
static unsigned long long splitmix(unsigned long long& x)
{
   unsigned long long z = (x += 0x9e3779b97f4a7c15ULL);
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   return z ^ (z >> 31);
}

int surrogate_create(int nin, int hid, int nout, unsigned long long seed, Surrogate** out)
{
   if (nin < 1 || nin > kMlpMaxIn || nout < 1 || nout > kMlpMaxOut || hid != kMlpHidden) {
      set_error("surrogate_create: need 1 <= n_in <= 4, 1 <= n_out <= 4, n_hidden = 64");
      return -1;
   }
   Surrogate* S = new Surrogate();
   S->nin = nin;
   S->hid = hid;
   S->nout = nout;
   S->nparam = S->off_b3() + nout;
   using L = MlpFrags<kMlpHidden>;
   hipError_t e = hipMalloc(&S->w, sizeof(float) * S->nparam);
   if (e == hipSuccess) e = hipMalloc(&S->m, sizeof(float) * S->nparam);
   if (e == hipSuccess) e = hipMalloc(&S->v, sizeof(float) * S->nparam);
   if (e == hipSuccess) e = hipMalloc(&S->frag, sizeof(u16x8) * L::NALL * 64);
   if (e == hipSuccess) e = hipMalloc(&S->bias, sizeof(float) * L::NBIAS);
   if (e != hipSuccess) {
      set_error(std::string("surrogate_create: ") + hipGetErrorString(e));
      surrogate_destroy(S);
      return -1;
   }
   // Glorot-uniform weights, zero biases
   std::vector<float> w(S->nparam, 0.0f);
   unsigned long long st = seed;
   auto fill = [&](long long off, int fan_out, int fan_in) {
      double a = std::sqrt(6.0 / (fan_in + fan_out));
      for (long long i = 0; i < (long long)fan_out * fan_in; i++) {
         double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
         w[off + i] = (float)((2.0 * u - 1.0) * a);
      }
   };
   fill(0, hid, nin);
   fill(S->off_w2(), hid, hid);
   fill(S->off_w3(), nout, hid);
   if (surrogate_set_params(S, w.data())) {
      surrogate_destroy(S);
      return -1;
   }
   *out = S;
   return 0;
}

void surrogate_destroy(Surrogate* S)
{
   if (!S) return;
   (void)hipFree(S->w);
   (void)hipFree(S->m);
   (void)hipFree(S->v);
   (void)hipFree(S->frag);
   (void)hipFree(S->bias);
   delete S;
}

int surrogate_set_params(Surrogate* S, const float* host)
{
   AMPE_HIP_CHECK(hipMemcpy(S->w, host, sizeof(float) * S->nparam, hipMemcpyHostToDevice));
   AMPE_HIP_CHECK(hipMemset(S->m, 0, sizeof(float) * S->nparam));
   AMPE_HIP_CHECK(hipMemset(S->v, 0, sizeof(float) * S->nparam));
   S->adam_t = 0;
   if (pack(S, nullptr)) return -1;
   AMPE_HIP_CHECK(hipDeviceSynchronize());
   return 0;
}

int surrogate_get_params(const Surrogate* S, float* host)
{
   AMPE_HIP_CHECK(hipMemcpy(host, S->w, sizeof(float) * S->nparam, hipMemcpyDeviceToHost));
   return 0;
}

int surrogate_forward(const Surrogate* S, long long n, const float* x, float* y, hipStream_t st)
{
   if (n <= 0) return 0;
   long long nb = (n + 255) / 256;
   hipLaunchKernelGGL(k_mlp_forward<kMlpHidden>, dim3((unsigned)nb), dim3(256), 0, st, S->dev(),
                      n, x, y);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

namespace {
struct GradPlan {
   int tpw;
   long long waves, blocks;
   int ncol;
   size_t lds;
};
GradPlan grad_plan(const Surrogate* S, long long n)
{
   using L = MlpFrags<kMlpHidden>;
   GradPlan g;
   long long tiles = (n + 15) / 16;
   g.tpw = (int)((tiles + 511) / 512);
   if (g.tpw < 1) g.tpw = 1;
   g.waves = (tiles + g.tpw - 1) / g.tpw;
   g.blocks = (g.waves + 3) / 4;
   g.waves = g.blocks * 4;
   g.ncol = (int)S->nparam + 1;
   g.lds = sizeof(u16x8) * L::NALL * 64 + sizeof(float) * (L::NBIAS + 4 * 16 * 17);
   return g;
}
}  // namespace

static int grad_launch(const Surrogate* S, const GradPlan& g, long long n, long long base,
                       long long ld, const float* x, const float* t, float* partial,
                       float* grad, double* dloss, hipStream_t st)
{
   hipLaunchKernelGGL(k_mlp_grad<kMlpHidden>, dim3((unsigned)g.blocks), dim3(256), g.lds, st,
                      S->dev(), n, base, ld, g.tpw, x, t, (float)(1.0 / (double)n), partial,
                      g.ncol);
   AMPE_HIP_CHECK(hipGetLastError());
   hipLaunchKernelGGL(k_mlp_reduce, dim3((g.ncol + 63) / 64), dim3(64 * kRedWaves), 0, st,
                      partial, (int)g.waves, g.ncol, 1.0 / (double)n, grad, dloss);
   AMPE_HIP_CHECK(hipGetLastError());
   return 0;
}

int surrogate_loss_grad(Surrogate* S, long long n, const float* x, const float* t, float* grad,
                        double* loss, hipStream_t st)
{
   if (n <= 0) {
      set_error("surrogate_loss_grad: n must be positive");
      return -1;
   }
   GradPlan g = grad_plan(S, n);
   float* partial = nullptr;
   double* dl = nullptr;
   AMPE_HIP_CHECK(hipMalloc(&partial, sizeof(float) * g.waves * g.ncol));
   if (hipMalloc(&dl, sizeof(double)) != hipSuccess) {
      (void)hipFree(partial);
      set_error("surrogate_loss_grad: out of device memory");
      return -1;
   }
   int rc = grad_launch(S, g, n, 0, n, x, t, partial, grad, dl, st);
   double h = 0.0;
   if (!rc && (hipMemcpyAsync(&h, dl, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
               hipStreamSynchronize(st) != hipSuccess)) {
      set_error("surrogate_loss_grad: copy failed");
      rc = -1;
   }
   (void)hipFree(partial);
   (void)hipFree(dl);
   if (loss) *loss = h;
   return rc;
}

int surrogate_train(Surrogate* S, long long n, const double* x, const double* t,
                    const ampe_train_params& tp, float* loss_hist, hipStream_t st)
{
   if (n <= 0 || tp.batch <= 0 || tp.batch > n || tp.steps <= 0) {
      set_error("surrogate_train: need 0 < batch <= n and steps > 0");
      return -1;
   }
   GradPlan g = grad_plan(S, tp.batch);
   float *xn = nullptr, *tn = nullptr, *partial = nullptr, *grad = nullptr;
   double* dl = nullptr;
   hipError_t e = hipMalloc(&xn, sizeof(float) * S->nin * n);
   if (e == hipSuccess) e = hipMalloc(&tn, sizeof(float) * S->nout * n);
   if (e == hipSuccess) e = hipMalloc(&partial, sizeof(float) * g.waves * g.ncol);
   if (e == hipSuccess) e = hipMalloc(&grad, sizeof(float) * S->nparam);
   if (e == hipSuccess) e = hipMalloc(&dl, sizeof(double) * tp.steps);
   int rc = 0;
   if (e != hipSuccess) {
      set_error(std::string("surrogate_train: ") + hipGetErrorString(e));
      rc = -1;
   }
   MlpDev m = S->dev();
   long long nb = (n + 255) / 256;
   if (!rc) {
      hipLaunchKernelGGL(k_mlp_norm, dim3((unsigned)nb), dim3(256), 0, st, n, S->nin, 0, x, m, xn);
      hipLaunchKernelGGL(k_mlp_norm, dim3((unsigned)nb), dim3(256), 0, st, n, S->nout, 1, t, m,
                         tn);
      if (hipGetLastError() != hipSuccess) {
         set_error("surrogate_train: normalisation launch failed");
         rc = -1;
      }
   }
   long long nbatch = n / tp.batch;
   unsigned pb = (unsigned)((S->nparam + 255) / 256);
   for (int k = 0; !rc && k < tp.steps; k++) {
      long long base = (k % nbatch) * tp.batch;
      rc = grad_launch(S, g, tp.batch, base, n, xn, tn, partial, grad, dl + k, st);
      if (rc) break;
      S->adam_t++;
      double frac = tp.steps > 1 ? (double)k / (tp.steps - 1) : 0.0;
      double lr = tp.lr0 * std::pow(tp.lr1 / tp.lr0, frac);
      float bc1 = (float)(1.0 - std::pow(tp.beta1, S->adam_t));
      float bc2 = (float)(1.0 - std::pow(tp.beta2, S->adam_t));
      hipLaunchKernelGGL(k_mlp_adam, dim3(pb), dim3(256), 0, st, S->nparam, S->w, S->m, S->v,
                         grad, (float)lr, (float)tp.beta1, (float)tp.beta2, (float)tp.eps, bc1,
                         bc2);
      if (hipGetLastError() != hipSuccess) {
         set_error("surrogate_train: Adam launch failed");
         rc = -1;
         break;
      }
      rc = pack(S, st);
   }
   if (!rc && loss_hist) {
      std::vector<double> h(tp.steps);
      if (hipMemcpyAsync(h.data(), dl, sizeof(double) * tp.steps, hipMemcpyDeviceToHost, st) !=
          hipSuccess) {
         set_error("surrogate_train: copy failed");
         rc = -1;
      }
      if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = -1;
      for (int k = 0; !rc && k < tp.steps; k++) loss_hist[k] = (float)h[k];
   }
   if (!loss_hist && hipStreamSynchronize(st) != hipSuccess) rc = -1;
   (void)hipFree(xn);
   (void)hipFree(tn);
   (void)hipFree(partial);
   (void)hipFree(grad);
   (void)hipFree(dl);
   return rc;
}

}  // namespace ampe
