/*
 * Oracle for the KKS surrogate path and the per-cell-temperature KKS batch.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP surrogate kernels and of
 * ampe_kks_batch; never linked into the product library.
 *
 * - o_calphad_gfun / o_calphad_at: end-member G(T) of the
 *   thermodynamic_data/calphadAuNi.dat:5-45 form and Redlich-Kister
 *   L_i(T) = L_i0 + L_i1 T (calphadAuNi.dat:47-63).  Pinned in the tests
 *   against the database values transcribed from that file.
 * - o_kks_newton / o_kks_batch: the binary CALPHAD KKS Newton of
 *   o_kks_cell (per-cell loop of CALPHADequilibriumPhaseConcentrationsStrategy::
 *   computePhaseConcentrationsOnPatch, source/
 *   CALPHADequilibriumPhaseConcentrationsStrategy.cc:343-402) with the end
 *   members evaluated at each cell's temperature, default start (0.5, 0.5)
 *   (tests/testCALPHADbinaryKKS.cc:123-127).  Pinned by the chemical
 *   potential equality and mass balance of every converged cell.
 * - o_mlp_forward / o_mlp_loss_grad: FP64 forward pass and backpropagation of
 *   the n_in -> H -> H -> n_out tanh MLP (no reference counterpart; the
 *   surrogate replaces the initial guess of
 *   CALPHADequilibriumPhaseConcentrationsStrategy.cc:368-369).  With
 *   emulate = 1 every matrix-core operand (weights, activations, upstream
 *   gradients) is rounded to BF16 as the GPU kernels round it and the hidden
 *   activations are stored in FP32; with emulate = 0 the arithmetic is plain
 *   FP64, and the tests pin that gradient against central finite differences
 *   of the loss.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define R_GAS 8.314472

float o_bf16(float f)
{
   uint32_t u;
   memcpy(&u, &f, 4);
   u += 0x7fffu + ((u >> 16) & 1u);
   u &= 0xffff0000u;
   float r;
   memcpy(&r, &u, 4);
   return r;
}

double o_calphad_gfun(const ampe_calphad_gfun* f, double T)
{
   int k = 0;
   while (k < f->nrange - 1 && T >= f->Tc[k + 1]) k++;
   const double* a = f->coef[k];
   double T2 = T * T, T3 = T2 * T;
   double T7 = T3 * T3 * T;
   double Tm9 = 1.0 / (T3 * T3 * T3);
   double g = a[0] + a[1] * T + a[2] * T * log(T);
   g = g + a[3] * T2;
   g = g + a[4] * T3;
   g = g + a[5] * T7;
   g = g + a[6] / T;
   g = g + a[7] * Tm9;
   return g;
}

void o_calphad_at(const ampe_calphad_tdep* tab, double T, ampe_calphad_phase* out)
{
   for (int ph = 0; ph < 2; ph++) {
      out[ph].gA = o_calphad_gfun(&tab->g[ph][0], T);
      out[ph].gB = o_calphad_gfun(&tab->g[ph][1], T);
      for (int i = 0; i < 4; i++) out[ph].L[i] = tab->L[ph][i][0] + tab->L[ph][i][1] * T;
   }
}

static double mu(const ampe_calphad_phase* cp, double T, double c)
{
   double RT = R_GAS * T;
   double x = 2.0 * c - 1.0;
   double P = cp->L[0] + x * (cp->L[1] + x * (cp->L[2] + x * cp->L[3]));
   double dP = cp->L[1] + x * (2.0 * cp->L[2] + x * 3.0 * cp->L[3]);
   return cp->gA - cp->gB + RT * log(c / (1.0 - c)) + (1.0 - 2.0 * c) * P +
          2.0 * c * (1.0 - c) * dP;
}

static double d2(const ampe_calphad_phase* cp, double T, double c)
{
   double RT = R_GAS * T;
   double x = 2.0 * c - 1.0;
   double P = cp->L[0] + x * (cp->L[1] + x * (cp->L[2] + x * cp->L[3]));
   double dP = cp->L[1] + x * (2.0 * cp->L[2] + x * 3.0 * cp->L[3]);
   double d2P = 2.0 * cp->L[2] + x * 6.0 * cp->L[3];
   return RT / (c * (1.0 - c)) - 2.0 * P + 4.0 * (1.0 - 2.0 * c) * dP +
          4.0 * c * (1.0 - c) * d2P;
}

/* Newton on F = [(mu_L - mu_S)/RT, (1-h) cL + h cS - c]; returns updates,
 * -1 without convergence, -2 when the residual after probe_its updates
 * exceeds probe_tol (probe_its < 0: no probe). */
int o_kks_newton(const ampe_calphad_phase* ph, double T, double h, double c, double* cl,
                 double* cs, int maxits, double tol, int probe_its, double probe_tol)
{
   double RTinv = 1.0 / (R_GAS * T);
   double xl = *cl, xs = *cs;
   const double lo = 1.0e-12, hi = 1.0 - 1.0e-12;
   if (!(xl > lo)) xl = lo;
   if (!(xl < hi)) xl = hi;
   if (!(xs > lo)) xs = lo;
   if (!(xs < hi)) xs = hi;
   for (int it = 0; it <= maxits; it++) {
      double f1 = (mu(&ph[0], T, xl) - mu(&ph[1], T, xs)) * RTinv;
      double f2 = (1.0 - h) * xl + h * xs - c;
      if (fabs(f1) < tol && fabs(f2) < tol) {
         *cl = xl;
         *cs = xs;
         return it;
      }
      if (it == maxits) break;
      if (it == probe_its && !(fmax(fabs(f1), fabs(f2)) <= probe_tol)) {
         *cl = xl;
         *cs = xs;
         return -2;
      }
      double a = d2(&ph[0], T, xl) * RTinv;
      double b = -d2(&ph[1], T, xs) * RTinv;
      double cc = 1.0 - h, d = h;
      double det = a * d - b * cc;
      double dl = -(d * f1 - b * f2) / det;
      double ds = -(a * f2 - cc * f1) / det;
      double alpha = 1.0;
      for (int k = 0; k < 60; k++) {
         double nl = xl + alpha * dl, ns = xs + alpha * ds;
         if (nl > 0.0 && nl < 1.0 && ns > 0.0 && ns < 1.0) break;
         alpha *= 0.5;
      }
      xl = xl + alpha * dl;
      xs = xs + alpha * ds;
   }
   *cl = xl;
   *cs = xs;
   return -1;
}

/* Batch KKS from the default start (0.5, 0.5), or from (gl, gs) when given
 * (with the probe / restart rule of ampe_kks_batch).  Returns the number of
 * non-converged cells; iters[i] = updates or -1. */
long o_kks_batch(const ampe_model_params* p, const ampe_calphad_tdep* tab, long n,
                 const double* c, const double* phi, const double* T, const double* gl,
                 const double* gs, int probe_its, double probe_tol, double* cl, double* cs,
                 int* iters)
{
   long fails = 0;
#pragma omp parallel for reduction(+ : fails)
   for (long i = 0; i < n; i++) {
      ampe_calphad_phase ph[2];
      o_calphad_at(tab, T[i], ph);
      double h = o_interp_func(phi[i], p->conc_interp);
      double xl = gl ? gl[i] : 0.5, xs = gs ? gs[i] : 0.5;
      int r = o_kks_newton(ph, T[i], h, c[i], &xl, &xs, p->newton_max_its, p->newton_tol,
                           gl ? probe_its : -1, probe_tol);
      if (r == -2) {
         xl = 0.5;
         xs = 0.5;
         int r2 = o_kks_newton(ph, T[i], h, c[i], &xl, &xs, p->newton_max_its, p->newton_tol,
                               -1, 0.0);
         r = r2 < 0 ? -1 : r2 + probe_its;
      }
      cl[i] = xl;
      cs[i] = xs;
      iters[i] = r;
      if (r < 0) fails++;
   }
   return fails;
}

/* ---------------------------------------------------------------- MLP */
typedef struct {
   int nin, hid, nout;
   long ob1, ow2, ob2, ow3, ob3, np;
} Lay;

static Lay lay(int nin, int hid, int nout)
{
   Lay L;
   L.nin = nin;
   L.hid = hid;
   L.nout = nout;
   L.ob1 = (long)hid * nin;
   L.ow2 = L.ob1 + hid;
   L.ob2 = L.ow2 + (long)hid * hid;
   L.ow3 = L.ob2 + hid;
   L.ob3 = L.ow3 + (long)nout * hid;
   L.np = L.ob3 + nout;
   return L;
}

static double rw(float w, int emulate) { return emulate ? (double)o_bf16(w) : (double)w; }
static double ra(double a, int emulate) { return emulate ? (double)o_bf16((float)a) : a; }

/* One sample: activations h1, h2 (hid), outputs y (nout).  In emulate mode
 * h1, h2 are rounded to FP32 (the GPU keeps them in FP32 registers). */
static void fwd1(const Lay* L, const float* w, const double* x, double* h1, double* h2,
                 double* y, int emulate)
{
   int H = L->hid;
   for (int u = 0; u < H; u++) {
      double z = w[L->ob1 + u];
      for (int i = 0; i < L->nin; i++) z += rw(w[(long)u * L->nin + i], emulate) * ra(x[i], emulate);
      h1[u] = emulate ? (double)(float)tanh(z) : tanh(z);
   }
   for (int u = 0; u < H; u++) {
      double z = w[L->ob2 + u];
      for (int k = 0; k < H; k++) z += rw(w[L->ow2 + (long)u * H + k], emulate) * ra(h1[k], emulate);
      h2[u] = emulate ? (double)(float)tanh(z) : tanh(z);
   }
   for (int o = 0; o < L->nout; o++) {
      double z = w[L->ob3 + o];
      for (int k = 0; k < H; k++) z += rw(w[L->ow3 + (long)o * H + k], emulate) * ra(h2[k], emulate);
      y[o] = z;
   }
}

/* x [nin][n], y [nout][n] (normalised data) */
void o_mlp_forward(int nin, int hid, int nout, const float* w, long n, const float* x,
                   double* y, int emulate)
{
   Lay L = lay(nin, hid, nout);
#pragma omp parallel for
   for (long s = 0; s < n; s++) {
      double xs[4], h1[256], h2[256], ys[4];
      for (int i = 0; i < nin; i++) xs[i] = x[i * n + s];
      fwd1(&L, w, xs, h1, h2, ys, emulate);
      for (int o = 0; o < nout; o++) y[o * n + s] = ys[o];
   }
}

/* Mean-squared error (1/n) sum (y - t)^2 and its gradient (grad[np]). */
double o_mlp_loss_grad(int nin, int hid, int nout, const float* w, long n, const float* x,
                       const float* t, double* grad, int emulate)
{
   Lay L = lay(nin, hid, nout);
   int H = hid;
   memset(grad, 0, sizeof(double) * L.np);
   double loss = 0.0;
   for (long s = 0; s < n; s++) {
      double xs[4], h1[256], h2[256], y[4], dy[4], dz2[256], dz1[256];
      for (int i = 0; i < nin; i++) xs[i] = x[i * n + s];
      fwd1(&L, w, xs, h1, h2, y, emulate);
      for (int o = 0; o < nout; o++) {
         double e = y[o] - (double)t[o * n + s];
         loss += e * e;
         dy[o] = 2.0 * e / (double)n;
      }
      for (int o = 0; o < nout; o++) {
         grad[L.ob3 + o] += dy[o];
         for (int k = 0; k < H; k++) grad[L.ow3 + (long)o * H + k] += ra(dy[o], emulate) * ra(h2[k], emulate);
      }
      for (int k = 0; k < H; k++) {
         double a = 0.0;
         for (int o = 0; o < nout; o++) a += rw(w[L.ow3 + (long)o * H + k], emulate) * ra(dy[o], emulate);
         dz2[k] = a * (1.0 - h2[k] * h2[k]);
         grad[L.ob2 + k] += dz2[k];
      }
      for (int u = 0; u < H; u++)
         for (int k = 0; k < H; k++) grad[L.ow2 + (long)u * H + k] += ra(dz2[u], emulate) * ra(h1[k], emulate);
      for (int k = 0; k < H; k++) {
         double a = 0.0;
         for (int u = 0; u < H; u++) a += rw(w[L.ow2 + (long)u * H + k], emulate) * ra(dz2[u], emulate);
         dz1[k] = a * (1.0 - h1[k] * h1[k]);
         grad[L.ob1 + k] += dz1[k];
      }
      for (int u = 0; u < H; u++)
         for (int i = 0; i < nin; i++) grad[(long)u * nin + i] += ra(dz1[u], emulate) * ra(xs[i], emulate);
   }
   return loss / (double)n;
}
