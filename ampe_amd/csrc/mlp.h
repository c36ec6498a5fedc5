// KKS surrogate MLP: n_in -> H -> H -> n_out, tanh hidden layers, on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulation).
//
// Matrix layout.  A layer is computed transposed, Z^T = W H^T + b, with the
// weights as the A operand (16 output rows x 32 K slots per instruction) and
// the activations of 16 samples as the B operand (32 K slots x 16 samples).
// v_mfma_f32_16x16x32_bf16 lane l holds A row l%16 / B column l%16 and K slots
// 8(l/16)..8(l/16)+7; the result lane l holds rows 4(l/16)..4(l/16)+3 of
// column l%16.  Two result tiles (rows 0-15 and 16-31 of a 32-feature block)
// therefore form the next layer's B operand without moving data between
// lanes if K slot s of the block stands for feature
//    perm(s) = 16 ((s%8)/4) + 4 (s/8) + s%4,
// and the packed weight fragments use the same permutation.  Backward data
// products (dH^T = W^T dZ^T) use transposed weight fragments the same way.
// Weight gradients (sums over samples) use v_mfma_f32_16x16x16_bf16 with the
// sample index as K, after a 16x16 transpose of each tile through LDS.
#pragma once

#include "physics.h"

namespace ampe {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpMaxIn = 4, kMlpMaxOut = 4, kMlpHidden = 64;

// Fragment image: one fragment = 64 lanes x 8 bf16 (one A operand).
template <int H>
struct MlpFrags {
   static constexpr int MT = H / 16, KB = H / 32;
   static constexpr int F1 = 0;              // W1 [mt]
   static constexpr int F2 = F1 + MT;        // W2 [mt][kb]
   static constexpr int F3 = F2 + MT * KB;   // W3 [kb]
   static constexpr int G2 = F3 + KB;        // W2^T [mt][kb]
   static constexpr int G3 = G2 + MT * KB;   // W3^T [mt]
   static constexpr int NFWD = G2;           // forward fragments
   static constexpr int NALL = G3 + MT;      // forward + backward
   static constexpr int NBIAS = 2 * H + 16;  // b1, b2, b3 (padded to 16)
};

// Device view of a surrogate.
struct MlpDev {
   int nin, nout;
   const u16x8* frag;  // fragment image, MlpFrags<H>::NALL * 64 entries
   const float* bias;  // NBIAS floats
   float in_shift[kMlpMaxIn], in_scale[kMlpMaxIn];
   double out_shift[kMlpMaxOut], out_scale[kMlpMaxOut];
};

__host__ __device__ inline int mlp_perm(int s)
{
   return 16 * ((s & 7) >> 2) + 4 * (s >> 3) + (s & 3);
}

// FP32 -> BF16, round to nearest even (finite inputs).
__host__ __device__ inline unsigned short f2bf(float f)
{
   unsigned u;
   __builtin_memcpy(&u, &f, 4);
   u += 0x7fffu + ((u >> 16) & 1u);
   return (unsigned short)(u >> 16);
}

// tanh(x) = 1 - 2 / (1 + e^{2x}); saturates correctly for large |x|.
__device__ inline float mlp_tanh(float x)
{
   float e = __expf(2.0f * x);
   return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ inline bf16x8 as_bf(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ inline bf16x8 pack8(f32x4 a, f32x4 b)
{
   u16x8 u;
   u[0] = f2bf(a[0]);
   u[1] = f2bf(a[1]);
   u[2] = f2bf(a[2]);
   u[3] = f2bf(a[3]);
   u[4] = f2bf(b[0]);
   u[5] = f2bf(b[1]);
   u[6] = f2bf(b[2]);
   u[7] = f2bf(b[3]);
   return as_bf(u);
}

__device__ inline f32x4 mfma32(u16x8 a, bf16x8 b, f32x4 c)
{
   return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a), b, c, 0, 0, 0);
}

// Copy the forward (nfrag = NFWD) or all fragments and the biases into LDS
// with the whole workgroup; the caller synchronises.
template <int H>
__device__ inline void mlp_stage(const MlpDev& m, int nfrag, u16x8* sf, float* sb)
{
   int nt = blockDim.x * blockDim.y;
   int t = threadIdx.x + blockDim.x * threadIdx.y;
   for (int i = t; i < nfrag * 64; i += nt) sf[i] = m.frag[i];
   for (int i = t; i < MlpFrags<H>::NBIAS; i += nt) sb[i] = m.bias[i];
}

// Forward pass of one 16-sample tile.  xb: layer-1 B operand (lanes 0-15
// hold the inputs of sample l%16 in K slots 0..nin-1).  h1, h2 receive the
// hidden activations (result layout), y the outputs (lanes 0-15, rows 0-3).
template <int H>
__device__ inline void mlp_tile_forward(const u16x8* __restrict__ F, const float* __restrict__ B,
                                        bf16x8 xb, f32x4* h1, f32x4* h2, f32x4& y)
{
   using L = MlpFrags<H>;
   const int lane = threadIdx.x & 63, g = lane >> 4;
#pragma unroll
   for (int mt = 0; mt < L::MT; mt++) {
      f32x4 acc;
      for (int i = 0; i < 4; i++) acc[i] = B[16 * mt + 4 * g + i];
      acc = mfma32(F[(L::F1 + mt) * 64 + lane], xb, acc);
      for (int i = 0; i < 4; i++) h1[mt][i] = mlp_tanh(acc[i]);
   }
   bf16x8 hb[L::KB];
#pragma unroll
   for (int kb = 0; kb < L::KB; kb++) hb[kb] = pack8(h1[2 * kb], h1[2 * kb + 1]);
#pragma unroll
   for (int mt = 0; mt < L::MT; mt++) {
      f32x4 acc;
      for (int i = 0; i < 4; i++) acc[i] = B[H + 16 * mt + 4 * g + i];
#pragma unroll
      for (int kb = 0; kb < L::KB; kb++)
         acc = mfma32(F[(L::F2 + mt * L::KB + kb) * 64 + lane], hb[kb], acc);
      for (int i = 0; i < 4; i++) h2[mt][i] = mlp_tanh(acc[i]);
   }
#pragma unroll
   for (int kb = 0; kb < L::KB; kb++) hb[kb] = pack8(h2[2 * kb], h2[2 * kb + 1]);
   for (int i = 0; i < 4; i++) y[i] = B[2 * H + 4 * g + i];
#pragma unroll
   for (int kb = 0; kb < L::KB; kb++) y = mfma32(F[(L::F3 + kb) * 64 + lane], hb[kb], y);
}

// Layer-1 B operand of tile nt of the wave: sample 16 nt + l%16 (its inputs
// live in that lane's x[]); lanes 16-63 carry zero K slots.
__device__ inline bf16x8 mlp_input_tile(const float* x, int nin, int nt)
{
   const int lane = threadIdx.x & 63;
   u16x8 u;
   for (int i = 0; i < 8; i++) {
      float v = __shfl(x[i < kMlpMaxIn ? i : 0], 16 * nt + (lane & 15));
      u[i] = (lane < 16 && i < nin) ? f2bf(v) : (unsigned short)0;
   }
   return as_bf(u);
}

// Whole-wave inference: every lane passes its own sample's normalised inputs
// x[0..nin) and receives y[0..nout).  All 64 lanes must call it.
template <int H>
__device__ inline void mlp_wave_infer(const u16x8* __restrict__ F, const float* __restrict__ B,
                                      int nin, const float* x, float* yout)
{
   using L = MlpFrags<H>;
   const int lane = threadIdx.x & 63;
#pragma unroll 1
   for (int nt = 0; nt < 4; nt++) {
      f32x4 h1[L::MT], h2[L::MT], y;
      mlp_tile_forward<H>(F, B, mlp_input_tile(x, nin, nt), h1, h2, y);
      for (int o = 0; o < kMlpMaxOut; o++) {
         float v = __shfl(y[o], lane & 15);
         if ((lane >> 4) == nt) yout[o] = v;
      }
   }
}

}  // namespace ampe

namespace ampe {

// Host object of a surrogate (device buffers).
struct Surrogate {
   int nin = 0, hid = 0, nout = 0;
   long long nparam = 0;
   float* w = nullptr;  // FP32 master parameters
   float* m = nullptr;  // Adam moments
   float* v = nullptr;
   u16x8* frag = nullptr;
   float* bias = nullptr;
   int adam_t = 0;
   float in_shift[kMlpMaxIn] = {0, 0, 0, 0}, in_scale[kMlpMaxIn] = {1, 1, 1, 1};
   double out_shift[kMlpMaxOut] = {0, 0, 0, 0}, out_scale[kMlpMaxOut] = {1, 1, 1, 1};
   long long off_b1() const { return (long long)hid * nin; }
   long long off_w2() const { return off_b1() + hid; }
   long long off_b2() const { return off_w2() + (long long)hid * hid; }
   long long off_w3() const { return off_b2() + hid; }
   long long off_b3() const { return off_w3() + (long long)nout * hid; }
   MlpDev dev() const;
};

int surrogate_create(int nin, int hid, int nout, unsigned long long seed, Surrogate** out);
void surrogate_destroy(Surrogate* s);
int surrogate_set_params(Surrogate* s, const float* host);
int surrogate_get_params(const Surrogate* s, float* host);
int surrogate_forward(const Surrogate* s, long long n, const float* x, float* y, hipStream_t st);
int surrogate_loss_grad(Surrogate* s, long long n, const float* x, const float* t, float* grad,
                        double* loss, hipStream_t st);
int surrogate_train(Surrogate* s, long long n, const double* x, const double* t,
                    const ampe_train_params& tp, float* loss_hist, hipStream_t st);

}  // namespace ampe
