// Device-resident BDF Newton-Krylov integrator (see sim.hip).
#pragma once

#include <type_traits>
#include <vector>

#include "amr.h"
#include "fac.h"
#include "kks.h"
#include "krylov.h"
#include "mg.h"

namespace ampe {

// Step and order state of the adaptive BDF (CVODE's fixed-leading-coefficient
// formulation on a Nordsieck history, orders 1..2): q the order, L = q + 1,
// qwait steps before an order change is considered, tau[i] the i-th previous
// step, l / tq the method and error-test coefficients of the current step,
// hscale the step the history is scaled to, eta / hprime / qprime the choice
// for the next step, etamax its growth bound.
struct BdfCtl {
   int q = 1, qprime = 1, L = 2, qwait = 2;
   double tau[4] = {0.0, 0.0, 0.0, 0.0};
   double l[3] = {1.0, 1.0, 0.0};
   double tq[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
   double hscale = 0.0, hprime = 0.0, eta = 1.0, etamax = 1.0e4, saved_tq5 = 0.0;
   bool started = false;
};

struct Sim {
   ampe_model_params p;
   ampe_solver_params s;
   Level* L = nullptr;
   Hier* hier = nullptr;  // patch hierarchy (L is then its composite level)
   double wsum = 0.0;  // summed control-volume weights of all components (hierarchy)
   int nc = 0, iq = -1, ic = -1;
   // phase orders (norderp > 1: MultiOrderRHSStrategy) and species (two for
   // the three-phase ternary model, tern)
   int nord = 1, ncs = 0;
   bool tern = false;
   double* ph6 = nullptr;  // ternary phase concentrations cL0, cL1, cA0, cA1, cB0, cB1
   double* Db = nullptr;   // solid-B face diffusivity (ternary EBS)
   long long ng = 0, nv = 0, nf = 0;
   double t = 0, h = 0, hprev = 0, gamma_setup = 1, crate = 1;
   long long nst = 0, nstlp = 0;
   bool have_prev = false;
   // state and Newton-Krylov work vectors (nc ghosted component blocks)
   double *y = nullptr, *yprev = nullptr, *ypred = nullptr, *ycur = nullptr,
          *ahist = nullptr, *fcur = nullptr, *b = nullptr, *del = nullptr, *ewt = nullptr,
          *work = nullptr, *work2 = nullptr, *jv = nullptr, *xk = nullptr, *corr = nullptr,
          *acor = nullptr;
   std::vector<double*> V;
   // adaptive mode: Nordsieck columns h y' and h^2 y''/2 (the state y is the
   // zeroth column); z2 holds the saved correction while the order is 1
   double *z1 = nullptr, *z2 = nullptr;
   BdfCtl bdf;
   std::vector<double> H;
   // phase concentrations and their per-step Newton guesses (ghosted)
   double *cl = nullptr, *cs = nullptr, *clr = nullptr, *csr = nullptr;
   // lagged face data and preconditioner coefficients
   double *n2 = nullptr, *Dc = nullptr, *Dphi = nullptr, *Dconc = nullptr, *fc = nullptr;
   double* fcq = nullptr;  // orientation face coefficients of the current RHS evaluation
   double *Cphi = nullptr, *sqrtm = nullptr, *Mphi = nullptr;
   double* phid = nullptr;  // phase field of the EBS isotropic stencil's diffusivities
   double *fac_u = nullptr, *fac_e = nullptr, *fac_r = nullptr;
   // FP32 preconditioner copies (mixed precision)
   float *f_u = nullptr, *f_e = nullptr, *f_r = nullptr, *f_f = nullptr;
   float *f_Cphi = nullptr, *f_sqrtm = nullptr, *f_Dconc = nullptr, *f_fc = nullptr,
         *f_Mphi = nullptr;
   hipStream_t stream = nullptr;
   hipEvent_t ev0 = nullptr, ev1 = nullptr;
   long long* kks_acc = nullptr;  // device KKS counters (launch_kks)
   KksGuess guess;                // surrogate guess when kks_guess is surrogate
   Mg<double>* mgd = nullptr;     // multigrid hierarchy (precond = MG, FP64)
   Mg<float>* mgf = nullptr;      // multigrid hierarchy (precond = MG, FP32)
   float last_ms = 0.f;
   ampe_stats st{};
   // captured FP32 preconditioner solve (graph_psolve) and the gamma its
   // kernel arguments hold
   hipGraphExec_t pgraph = nullptr;
   double pgraph_gamma = 0.0;
   // scalar temperature ramp (sim_set_temperature_ramp)
   bool tramp = false, has_tab = false;
   double T0 = 0.0, dTdt = 0.0, Ttarget = 0.0;
   ampe_calphad_tdep tab{};
};

int sim_create(const ampe_model_params& p, const ampe_solver_params& s, const ampe_box* boxes,
               int npatch, const ampe_box* periodic, Sim** out);
int sim_create_hier(const ampe_model_params& p, const ampe_solver_params& s, Hier* H, Sim** out);
void sim_destroy(Sim* S);
int sim_set_state_composite(Sim* S, const double* y_host);
int sim_get_state_composite(Sim* S, double* y_host);
int sim_set_time(Sim* S, double t, double h0);
int sim_set_state(Sim* S, const double* phi, const double* c, const double* q);
int sim_get_state(Sim* S, double* phi, double* c, double* q, double* cl, double* cs);
int sim_advance(Sim* S, int nsteps, ampe_stats* stats);
int sim_set_temperature_ramp(Sim* S, double T0, double dTdt, double T_target,
                             const ampe_calphad_tdep* tab);

int vec_mul_comp(Sim* S, double* x, const double* sm, int ncomp);

// Integrator callbacks on caller-owned state vectors (nv doubles, the sim's
// layout), enqueued on the sim's stream.
int sim_cb_rhs(Sim* S, double t, double* y, double* ydot, int fd);
int sim_cb_psetup(Sim* S, double t, double* y, double gamma);
int sim_cb_psolve(Sim* S, const double* r, double* z, double delta);
int sim_cb_project(Sim* S, const double* y, double* corr, double* err);
int sim_cb_reset_ref(Sim* S);
int sim_cb_kks_check(Sim* S);

}  // namespace ampe
