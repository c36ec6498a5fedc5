// Fused Newton-Krylov vector kernels (see krylov.hip).  Host results of the
// reductions are raw sums over interior cells (no cell-volume factor).
#pragma once

#include "kernels.h"

namespace ampe {

constexpr int kMaxKrylov = 32;

// FP32 preconditioner right-hand side written next to a Krylov vector r:
// f = r, orientation components [iq, iq + ql) divided by sqrt(M) = sm
// (f == nullptr: none).
struct PrecIn {
   float* f = nullptr;
   const double* sm = nullptr;
   int iq = 0, ql = 0;
   int split = 0;  // orientation components in the row parity-split layout
};

// A Krylov vector read from the FP32 preconditioner solution u: v = ewt u,
// orientation components multiplied by sm first (u == nullptr: read v).
struct PrecOut {
   const float* u = nullptr;
   const double* sm = nullptr;
   const double* ewt = nullptr;
   int iq = 0, ql = 0;
   int split = 0;  // orientation components in the row parity-split layout
};

struct VecList {
   int n;
   const double* v[kMaxKrylov];
   double y[kMaxKrylov];
};

int kry_prep(const Level& L, int nc, double hinv, double* V, const double* ewt, double* xk,
             double* sumsq, hipStream_t s);
int kry_mid(const Level& L, int nc, double sig, const double* xk, const double* ycur,
            double* work2, hipStream_t s);
// The same with the direction (V h) / ewt computed on the fly from an
// unnormalised basis vector V and its scale h (no xk array).
int kry_mid_scaled(const Level& L, int nc, double sig, const double* V, double h,
                   const double* ewt, const double* ycur, double* work2, hipStream_t s);
// (out may be null when pin.f receives the FP32 preconditioner input.)
int kry_finish_scaled(const Level& L, int nc, double siginv, double gamma, const double* V,
                      double h, const double* ewt, const double* jv, const double* fcur,
                      double* out, const PrecIn& pin, hipStream_t s);
int kry_finish(const Level& L, int nc, double siginv, double gamma, const double* xk,
               const double* jv, const double* fcur, double* out, hipStream_t s);
int kry_scale_dot(const Level& L, int nc, const double* ewt, const double* work, double* out,
                  const double* other, double* dot, hipStream_t s);
int kry_mgs(const Level& L, int nc, double* w, const double* Vi, double h, const double* Vnext,
            double* dot, hipStream_t s);
int kry_lincomb(const Level& L, int nc, const VecList& vl, const double* ewt, double* x,
                hipStream_t s);
int kry_newton_rhs(const Level& L, int nc, double gamma, const double* f, const double* y,
                   const double* ah, double* b, hipStream_t s);
// Classical Gram-Schmidt (at most kMaxCgs dots per pass, so maxl <= 14).
constexpr int kMaxCgs = 16;
// out = ewt * work (ewt != null) or out as is; sums[k] = out . vl.v[k]
// (k < vl.n) and sums[vl.n] = out . out, one pass and one synchronisation.
int kry_cgs_dots(const Level& L, int nc, const double* ewt, const double* work, double* out,
                 const VecList& vl, double* sums, hipStream_t s);
// The same with out = ewt u formed from the FP32 preconditioner solution.
int kry_cgs_dots_from(const Level& L, int nc, const PrecOut& src, double* out,
                      const VecList& vl, double* sums, hipStream_t s);
// w = w + sum_k vl.y[k] vl.v[k] (k in order); *sumsq = w . w.
int kry_cgs_update(const Level& L, int nc, double* w, const VecList& vl, double* sumsq,
                   hipStream_t s);
int kry_newton_update(const Level& L, int nc, const double* del, const double* ewt, double* y,
                      double* sumsq, hipStream_t s);

}  // namespace ampe
