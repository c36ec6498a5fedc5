/*
 * Oracle: a plain-C FP64 restatement of AMPE's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the HIP path is compared
 * against and the CPU baseline that bench.py times; it is never linked into,
 * called by or shipped with the product library (ampe_amd/).
 *
 * Every function cites the reference file:line it restates.  Arrays follow the
 * same layout as the product (one patch): a cell field is the patch box grown
 * by g ghost cells (no ghosts along k in 2-D), i fastest; a side field holds the
 * x faces ((nx+1) ny nz), then the y faces (nx (ny+1) nz), then the z faces.
 *
 * Parity status.  Each function below is checked by a named CPU test against
 * mathematics the function itself does not restate (closed forms, invariants,
 * numpy / scipy computations, the reference's unit-test vectors):
 *   functions.f helpers, free energies, KKS (quadratic, CALPHAD, dilute):
 *     tests/test_oracle.py, tests/test_dilute_oracle.py,
 *     tests/test_surrogate_oracle.py;
 *   simple / isotropic fluxes, quaternion gradients, composition kernels,
 *     RBGS / residual, computerhspbg (rotation symmetry): tests/test_oracle.py;
 *   anisotropic fluxes (2-D Kobayashi form, 3-D cubic form with a scipy
 *     rotation), orientation face coefficients (energy derivative),
 *     orientation mobility, orientation RHS (tangency, steady twist,
 *     continuum convergence, rotation equivariance), CVODE projection,
 *     tagging, dot / WRMS: tests/test_oracle_pins.py;
 *   EBS stencils: tests/test_ebs_oracle.py; multi-order RHS and projection:
 *     tests/test_multiorder_oracle.py; grain labels: tests/test_grains_oracle.py;
 *   integrator: BDF order of accuracy (tests/test_oracle.py) and the
 *     reference's regression decks end to end (tests/test_oracle_deck.py).
 *   parity unpinned: exact Thermo4PFM Newton iterate sequence; CVODE step
 *     heuristics (this is a BDF1/2 restatement); hypre PCG/PFMG replaced by
 *     red-black Gauss-Seidel on the same symmetrised matrix.
 */
#ifndef AMPE_ORACLE_H
#define AMPE_ORACLE_H

#include "../include/ampe_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OGeom {
   int ndim, nx, ny, nz, g, gz;
   long sy, sz, ncg;       /* strides and size of a ghosted cell array */
   long nfx, nfy, nfz;     /* face counts per direction */
   double dx[3];
   int periodic;           /* 1: periodic box; 0: zero-slope ghosts (clamped) */
} OGeom;

void o_geom_init(OGeom* G, int ndim, int nx, int ny, int nz, int g,
                 const double dx[3]);

/* functions.f helpers */
double o_interp_func(double phi, char type);
double o_deriv_interp_func(double phi, char type);
double o_second_deriv_interp_func(double phi, char type);
double o_well_func(double phi, char type);
double o_deriv_well_func(double phi, char type);
double o_second_deriv_well_func(double phi, char type);
double o_average_func(double a, double b, char type);
double o_eval_grad_normi(double gn2, char floor_type, double floor2,
                         double max_gni);

/* free energies */
double o_fenergy(const ampe_model_params* p, double c, int phase);
double o_dfenergy(const ampe_model_params* p, double c, int phase);
double o_d2fenergy(const ampe_model_params* p, double c, int phase);
int o_kks_cell(const ampe_model_params* p, double c, double phi, double* cl,
               double* cs);

int o_grains_label(const OGeom* G, const double* phase, double thr, int periodic, int* labels,
                   int* sweeps);
/* composition-dependent EBS diffusivities (CALPHAD mobilities) */
double o_atomic_mobility(const ampe_calphad_mobility* m, double c0, double c1, double T);
double o_phase_diffusion_mobility(const ampe_calphad_mobility m[2], double c0, double T);
void o_ebs_coeffs_mob(const OGeom* G, const ampe_model_params* p, const double* phi,
                      const double* cl, const double* cs, double* DL, double* DA);
void o_ebs_coeffs(const OGeom* G, const ampe_model_params* p, const double* phi, double* DL,
                  double* DA);
/* three phases on norderp > 1 orders, ternary quadratic (oracle_ternary.c) */
void o_tern3_phase_conc(const ampe_model_params* p, const double c[2], const double h[3],
                        double x[6]);
void o_kks_tern3(const OGeom* G, const ampe_model_params* p, int norder, const double* phi,
                 const double* c, double* ph);
double o_tern3_driving_force(const ampe_model_params* p, int norder, const double* phi_cell,
                             const double ph_cell[6], int ip);
void o_phase_rhs_tern3(const OGeom* G, const ampe_model_params* p, int norder,
                       const double* phi, const double* ph, const double* M, double* rhs);
void o_ebs3_coeffs(const OGeom* G, const ampe_model_params* p, int norder, const double* phi,
                   double* DL, double* DA, double* DB);
void o_conc_rhs_ebs3(const OGeom* G, const ampe_model_params* p, const double* ph,
                     const double* DL, const double* DA, const double* DB, double* rhs);
void o_phase_flux_multiorder(const OGeom* G, const ampe_model_params* p, int norder,
                             const double* phi, double* fl);
void o_phase_rhs_multiorder_fx(const OGeom* G, const ampe_model_params* p, int norder,
                               const double* fl, const double* phi, const double* M,
                               double* rhs);
void o_phase_rhs_tern3_fx(const OGeom* G, const ampe_model_params* p, int norder,
                          const double* fl, const double* phi, const double* ph,
                          const double* M, double* rhs);
void o_ebs3_flux(const OGeom* G, const ampe_model_params* p, const double* ph, const double* DL,
                 const double* DA, const double* DB, double* fl2);
void o_conc_rhs_ebs3_fx(const OGeom* G, const ampe_model_params* p, const double* fl2,
                        double* rhs);
void o_phase_rhs_multiorder(const OGeom* G, const ampe_model_params* p, int norder,
                            const double* phi, const double* M, double* rhs);
void o_project_phases(const OGeom* G, int norder, const double* phi, double* corr,
                      double* err);
void o_conc_rhs_ebs_4th(const OGeom* G, const ampe_model_params* p, const double* cl,
                        const double* cs, const double* DL, const double* DA, double* rhs);
void o_conc_rhs_ebs_iso(const OGeom* G, const ampe_model_params* p, const double* phid,
                        const double* cl, const double* cs, double* rhs);
void o_conc_rhs_ebs(const OGeom* G, const ampe_model_params* p, const double* cl,
                    const double* cs, const double* DL, const double* DA, double* rhs);
void o_ghost_fill_periodic(const OGeom* G, double* f, int ncomp);
/* ghosts from the periodic image (G->periodic) or from the nearest interior
 * cell, each index clamped: the zero-slope boundary of the reference's
 * "slope 0" conditions, as the device level's physical boundaries */
void o_ghost_fill(const OGeom* G, double* f, int ncomp);

void o_phase_flux(const OGeom* G, const ampe_model_params* p,
                  const double* phi, const double* q, double* flux);
void o_quat_grad_modulus(const OGeom* G, const ampe_model_params* p,
                         const double* q, double* gmod);
void o_quat_face_norm2(const OGeom* G, const ampe_model_params* p,
                       const double* q, double* n2);
void o_phase_rhs(const OGeom* G, const ampe_model_params* p,
                 const double* phi, const double* q, const double* cl,
                 const double* cs, double* rhs);
void o_phase_rhs_fx(const OGeom* G, const ampe_model_params* p, const double* flux,
                    const double* phi, const double* q, const double* cl,
                    const double* cs, double* rhs);
void o_quat_flux(const OGeom* G, const ampe_model_params* p, const double* phi,
                 const double* q, const double* n2, double* fl);
void o_quat_rhs_fx(const OGeom* G, const ampe_model_params* p, const double* phi,
                   const double* q, const double* fl, double* rhs);
void o_conc_flux(const OGeom* G, const ampe_model_params* p, const double* c,
                 const double* phi, const double* Dc, const double* Dphi, double* fl);
void o_conc_rhs_fx(const OGeom* G, const ampe_model_params* p, const double* fl, double* rhs);
void o_quat_mobility(const OGeom* G, const ampe_model_params* p,
                     const double* phi, double* mq);
void o_quat_face_coef(const OGeom* G, const ampe_model_params* p,
                      const double* phi, const double* n2, double* fc);
void o_quat_rhs(const OGeom* G, const ampe_model_params* p, const double* phi,
                const double* q, const double* n2, double* rhs);
void o_quat_project(const OGeom* G, const ampe_model_params* p,
                    const double* q, double* corr, double* err);
void o_conc_coeffs(const OGeom* G, const ampe_model_params* p,
                   const double* phi, const double* cl, const double* cs,
                   double* Dc, double* Dphi);
void o_conc_rhs(const OGeom* G, const ampe_model_params* p, const double* c,
                const double* phi, const double* Dc, const double* Dphi,
                double* rhs);
int o_kks_solve(const OGeom* G, const ampe_model_params* p, const double* phi,
                const double* c, const double* clr, const double* csr,
                double* cl, double* cs);
void o_tag_cells(const OGeom* G, const ampe_model_params* p,
                 const double* phi, double tol, int* tags);

/* elliptic FAC pieces */
void o_phase_fac_C(const OGeom* G, const ampe_model_params* p, double gamma,
                   const double* phi, const double* M, double* C);
double o_phase_mobility(const ampe_model_params* p, double cl, double cs);
void o_phase_mobility_field(const OGeom* G, const ampe_model_params* p, const double* cl,
                            const double* cs, double* M);
double o_rbgs_color(const OGeom* G, const ampe_elliptic_op* op,
                    const double* rhs, double* u, int color, double* flux);
double o_fac_smooth(const OGeom* G, const ampe_elliptic_op* op,
                    const double* rhs, double* u, int sweeps, double tol,
                    double* flux);
double o_fac_residual(const OGeom* G, const ampe_elliptic_op* op,
                      const double* rhs, double* u, double* res,
                      double* flux);

/* vector reductions (interior cells) */
double o_dot(const OGeom* G, int ncomp, const double* x, const double* y);
double o_wrms(const OGeom* G, int ncomp, const double* x, const double* w);

/* integrator */
typedef struct OSim OSim;
OSim* o_sim_create(const ampe_model_params* p, const ampe_solver_params* s,
                   int nx, int ny, int nz);
void o_sim_destroy(OSim* S);
void o_sim_set_state(OSim* S, const double* phi, const double* c,
                     const double* q);
void o_sim_get_state(OSim* S, double* phi, double* c, double* q, double* cl,
                     double* cs);
int o_sim_advance(OSim* S, int nsteps, ampe_stats* st);
/* physical (zero-slope) boundaries instead of the periodic box; call before
 * o_sim_set_state */
void o_sim_set_periodic(OSim* S, int periodic);
/* Scalar temperature T(t) = T0 + dTdt t, held at T_target once reached
 * (ScalarTemperatureStrategy::getCurrentTemperature,
 * source/ScalarTemperatureStrategy.cc), applied at the time of every step
 * attempt (evaluateRHSFunction sets the field at the evaluation time,
 * source/QuatIntegrator.cc:3191); with tab, the CALPHAD phases are
 * re-evaluated at that temperature. */
void o_sim_set_temperature_ramp(OSim* S, double T0, double dTdt, double T_target,
                                const ampe_calphad_tdep* tab);


/* surrogate path (oracle_surrogate.c) */
float o_bf16(float f);
double o_calphad_gfun(const ampe_calphad_gfun* f, double T);
void o_calphad_at(const ampe_calphad_tdep* tab, double T, ampe_calphad_phase* out);
int o_kks_newton(const ampe_calphad_phase* ph, double T, double h, double c, double* cl,
                 double* cs, int maxits, double tol, int probe_its, double probe_tol);
long o_kks_batch(const ampe_model_params* p, const ampe_calphad_tdep* tab, long n,
                 const double* c, const double* phi, const double* T, const double* gl,
                 const double* gs, int probe_its, double probe_tol, double* cl, double* cs,
                 int* iters);
void o_mlp_forward(int nin, int hid, int nout, const float* w, long n, const float* x,
                   double* y, int emulate);
double o_mlp_loss_grad(int nin, int hid, int nout, const float* w, long n, const float* x,
                       const float* t, double* grad, int emulate);

/* ---------------------------------------------------------------------------
 * Multi-level hierarchy (oracle_amr.c): levels of ratio 2, level 0 covering
 * the domain, ghost width 1; composite cell fields (per component, every
 * patch's ghosted block, level by level) and side fields (every patch's x, y,
 * z faces in a row).
 * ------------------------------------------------------------------------- */
typedef struct OHier OHier;
typedef struct OHBlock {
   int kind;                /* 0 elliptic (EllipticFACOps), 1 orientation block */
   const double* D;         /* side field or NULL */
   double D_const;
   const double* C;         /* cell field or NULL (elliptic) */
   double C_const;
   const double* M;         /* cell field or NULL (elliptic) */
   double M_const;
   const double* S;         /* orientation: sqrt(M_q), cells with ghosts */
   double gamma;            /* orientation: operator u + gamma S div(D grad(S u)) */
   int ewing;               /* Ewing's coarse-fine flux fix (elliptic blocks) */
} OHBlock;
typedef struct OHFacParams {
   int max_cycles, pre, post, coarse_its;
   double coarse_tol;
} OHFacParams;
OHier* o_hier_create(int ndim, int nlev, const int* npatch, const ampe_box* boxes,
                     const int dom0[3], const double dx0[3], int periodic);
void o_hier_destroy(OHier* H);
int o_hier_nlev(const OHier* H);
int o_hier_npatch(const OHier* H);
long o_hier_ncg(const OHier* H);
long o_hier_nf(const OHier* H);
void o_hier_offsets(const OHier* H, long* coff, long* foff);
const OGeom* o_hier_geom(const OHier* H, int p);
void o_hier_ghost_fill_level(const OHier* H, int l, double* f, int ncomp, int cf);
void o_hier_ghost_fill(const OHier* H, double* f, int ncomp, int cf);
void o_hier_weights(const OHier* H, double* w);
void o_hier_coarsen_level(const OHier* H, int l, double* f, int ncomp, int quatnorm);
void o_hier_coarsen(const OHier* H, double* f, int ncomp, int quatnorm);
int o_hier_transfer(const OHier* Ho, const OHier* Hn, const double* fo, double* fn, int ncomp);
void o_hier_coarsen_faces_to(const OHier* H, int l, const double* src, double* dst, int ncomp,
                             long stride);
void o_hier_coarsen_faces(const OHier* H, int l, double* s, int ncomp, long stride);
void o_hier_phase_rhs(const OHier* H, const ampe_model_params* p0, const double* phi,
                      const double* q, const double* cl, const double* cs, double* rhs);
void o_hier_quat_face_norm2(const OHier* H, const ampe_model_params* p0, const double* q,
                            double* n2);
void o_hier_quat_rhs(const OHier* H, const ampe_model_params* p0, const double* phi,
                     const double* q, const double* n2, double* rhs);
void o_hier_conc_coeffs(const OHier* H, const ampe_model_params* p0, const double* phi,
                        const double* cl, const double* cs, double* Dc, double* Dphi);
void o_hier_phase_rhs_tern3(const OHier* H, const ampe_model_params* p0, int norder,
                            const double* phi, const double* ph6, double* rhs);
void o_hier_phase_rhs_multiorder(const OHier* H, const ampe_model_params* p0, int norder,
                                 const double* phi, double* rhs);
void o_hier_ebs3_coeffs(const OHier* H, const ampe_model_params* p0, int norder,
                        const double* phi, double* DL, double* DA, double* DB);
void o_hier_conc_rhs_ebs3(const OHier* H, const ampe_model_params* p0, const double* ph6,
                          const double* DL, const double* DA, const double* DB, double* rhs);
void o_hier_conc_rhs(const OHier* H, const ampe_model_params* p0, const double* c,
                     const double* phi, const double* Dc, const double* Dphi, double* rhs);
double o_hier_comp_residual(const OHier* H, const OHBlock* B, double* u, const double* f,
                            double* r, int ncomp);
typedef struct OSim OSim;
OSim* o_sim_create_hier(const ampe_model_params* p, const ampe_solver_params* s, OHier* H);
void o_sim_set_state_composite(OSim* S, const double* y);
void o_sim_set_time(OSim* S, double t, double h0);
void o_sim_get_state_composite(OSim* S, double* y);
int o_hier_fac_solve(const OHier* H, const OHBlock* B, const OHFacParams* P, const double* f,
                     double* u, int ncomp, double tol, double* norm, int* cycles);

#ifdef __cplusplus
}
#endif
#endif
