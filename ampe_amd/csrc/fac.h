// FAC preconditioner launch interface (see fac.hip).
#pragma once

#include "kernels.h"

namespace ampe {

// A u = M div(D grad u) + C u; a NULL field pointer selects the constant.
template <typename T>
struct EllOp {
   const T* D = nullptr;
   T Dc = 0;
   const T* C = nullptr;
   T Cc = 0;
   const T* M = nullptr;
   T Mc = 0;
};

int launch_phase_C(const Level& L, const ampe_model_params& p, double gamma,
                   const double* phi, const double* M, double* C, hipStream_t s);

// One colour sweep; with want_max the per-tile max |res| lands in L.partial.
template <typename T>
int launch_rbgs(const Level& L, const EllOp<T>& op, const double* dx, const T* rhs, T* u,
                int color, bool want_max, hipStream_t s);

// res = rhs - A u; if part is not null the per-tile sum res^2 lands in
// part[0 .. L.ncell_tiles).
template <typename T>
int launch_residual(const Level& L, const EllOp<T>& op, const double* dx, const T* rhs,
                    const T* u, T* res, double* part, hipStream_t s);

// Orientation block: one colour of the red-black sweep, and the residual.
// split: the component blocks of rhs, z and res are stored in the row
// parity-split layout (split_col; fc and sm stay natural).
template <typename T>
int launch_quat_rbgs(const Level& L, const T* fc, const T* sm, double gamma, const double* dx,
                     int ncomp, const T* rhs, T* z, int color, hipStream_t s, bool split = false);

template <typename T>
int launch_quat_residual(const Level& L, const T* fc, const T* sm, double gamma,
                         const double* dx, int ncomp, const T* rhs, const T* z, T* res,
                         double* part, hipStream_t s, bool split = false);

// b (on coarse, the fine boxes halved) = the restriction of the elliptic
// residual rhs - A u on fine after red-black sweeps ending with the black
// colour: the red cells' residuals alone, summed per coarse cell (periodic
// levels with one checkerboard; the caller checks).
template <typename T>
int launch_ell_resrestrict(const Level& coarse, const Level& fine, const EllOp<T>& op,
                           const double* dx, const T* rhs, const T* u, T* b, hipStream_t s);

// b (on coarse, the fine boxes halved) = the restriction of the residual
// rhs - A z on fine when z comes out of red-black sweeps ending with the black
// colour: the red cells' residuals alone (the black ones vanish), summed per
// coarse cell.  One, two or four components; valid on periodic levels whose
// colouring is one checkerboard (the caller checks).
template <typename T>
int launch_quat_resrestrict(const Level& coarse, const Level& fine, const T* fc, const T* sm,
                            double gamma, const double* dx, int ncomp, const T* rhs, const T* z,
                            T* b, hipStream_t s, bool split = false);

// ------------------------------------------------------------ vectors
// Flat element-wise ops over n entries (whole ghosted vectors).
// One red-black sweep of the orientation block from zin (ghosts filled) into
// zout in one pass (plus a ghost fill of zout and a pass over the patches'
// outer layer): the values of the red launch followed by the black launch.
template <typename T>
int launch_quat_rbgs_fused(const Level& L, const T* fc, const T* sm, double gamma,
                           const double* dx, int ncomp, const T* rhs, const T* zin, T* zout,
                           hipStream_t s);
int vec_axpby(long long n, double a, const double* x, double b, const double* y, double* z,
              hipStream_t s);
int vec_fill(long long n, double v, double* z, hipStream_t s);
int vec_mul(long long n, const double* x, const double* y, double* z, hipStream_t s);
int vec_div(long long n, const double* x, const double* y, double* z, hipStream_t s);
int vec_ewt(long long n, const double* y, double rtol, double atol, double* w, hipStream_t s);
int vec_bdf2_hist(long long n, double c1, const double* y, double c2, const double* yp,
                  double* out, hipStream_t s);
int vec_to_f32(long long n, const double* x, float* y, hipStream_t s);
int vec_axpby_f(long long n, float* u, const float* e, hipStream_t s);  // u += e
// z[0, n) = 0 by a kernel (usable under HIP graph capture)
template <typename T>
int fill_zero(T* z, long long n, hipStream_t s);
int vec_to_f64(long long n, const float* x, double* y, hipStream_t s);
int vec_div_f(long long n, const double* x, const double* sm, double* y, hipStream_t s);
int vec_sqrt(long long n, const double* x, double* y, hipStream_t s);
// Preconditioner input / output over nc component blocks (orientation blocks
// [iq, iq+ql) scaled by 1/sqrt(M) on input, sqrt(M) on output); psolve_in
// also zeroes u.  With f null (FP64) only the orientation blocks go to fq.
// psolve_out multiplies the result by ewt when ewt is not null.  split (a
// level): the orientation blocks of f / fq / u are in that level's row
// parity-split layout.
template <typename T>
int psolve_in(long long ng, int nc, int iq, int ql, const double* r, const double* sm, T* f,
              T* fq, T* u, hipStream_t s, const Level* split = nullptr);
template <typename T>
int psolve_out(long long ng, int nc, int iq, int ql, const T* u, const double* sm,
               const double* ewt, double* z, hipStream_t s, const Level* split = nullptr);
int vec_scale_comp(long long ng, int ncomp, const double* sm, double* x, hipStream_t s);
// Reductions over interior cells of ncomp component blocks (host result,
// synchronises s).
int vec_dot(const Level& L, int ncomp, const double* x, const double* y, double* out,
            hipStream_t s);
int vec_sumsq_w(const Level& L, int ncomp, const double* x, const double* w, double* out,
                hipStream_t s);

}  // namespace ampe
