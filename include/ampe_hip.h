/*
 * ampe_hip.h -- C ABI of the MI355X (gfx950) implementation of AMPE's
 * per-timestep hot path: phase / orientation / composition right-hand sides,
 * per-cell KKS phase concentrations, the FAC (red-black Gauss-Seidel)
 * preconditioner blocks, CVODE-style vector reductions, AMR gradient tagging
 * and a BDF Newton-Krylov time integrator that drives them.
 *
 * Conventions (all entry points):
 *   - return 0 on success, >0 for a recoverable condition (e.g. a solver that
 *     did not converge), <0 for a fatal error; ampe_last_error() gives text.
 *     No entry point aborts the process.
 *   - "device" pointers are HIP device allocations owned by the caller;
 *     no entry point allocates on the hot path except the *_create calls.
 *   - stream is an opaque hipStream_t (NULL = default stream).  Entry points
 *     enqueue work and do not synchronise the device, except those that return
 *     a host scalar (reductions, KKS iteration counts), which synchronise the
 *     given stream.
 *
 * Data layout on a level (SAMRAI CellData/SideData semantics, i fastest):
 *   cell field   : [component][patch][ghost box]  ghost box = box grown by g
 *                  (in 2-D the k extent is 1 and carries no ghosts)
 *   side field   : [direction][patch][faces]      faces of direction d have
 *                  extent n_d + 1 along d and no ghosts
 *   patches of one level are stored back to back; per-patch offsets are
 *   available from ampe_level_patch_offsets().
 *
 * Each entry cites the reference interface it replaces (file:line in
 * LLNL/AMPE).
 */
#ifndef AMPE_HIP_H
#define AMPE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void* ampe_stream_t;

/* Cell-centred index box, inclusive bounds (SAMRAI hier::Box). */
typedef struct ampe_box {
   int lo[3];
   int hi[3];
} ampe_box;

enum ampe_flux_type {
   AMPE_FLUX_SIMPLE = 0,      /* PhaseFluxStrategySimple -> gradient_flux   */
   AMPE_FLUX_ISOTROPIC = 1,   /* PhaseFluxStrategyIsotropic (2-D only)      */
   AMPE_FLUX_ANISOTROPIC = 2  /* PhaseFluxStrategyAnisotropy                */
};

enum ampe_free_energy {
   AMPE_FE_QUADRATIC = 0,      /* QuadraticFreeEnergyStrategy              */
   AMPE_FE_CALPHAD_BINARY = 1, /* CALPHADFreeEnergyStrategyBinary          */
   AMPE_FE_KKS_DILUTE = 2,     /* KKSdiluteBinary (dilute binary solution) */
   AMPE_FE_QUADRATIC_TERNARY_3PH = 3 /* QuadraticFreeEnergyStrategyMultiOrderTernaryThreePhase:
                                      * three phases (liquid, solid A, solid B) on
                                      * norderp > 1 order parameters, two species */
};

/* Binary CALPHAD phase description evaluated at the (uniform) temperature:
 * G(c) = c*gA + (1-c)*gB + RT[c ln c + (1-c) ln(1-c)]
 *        + c(1-c) * sum_i L[i] (2c-1)^i                      [J/mol]
 * c is the mole fraction of species A (first species of the database). */
typedef struct ampe_calphad_phase {
   double gA, gB;
   double L[4];
} ampe_calphad_phase;

/* CALPHAD atomic mobility of one species in one phase (CALPHADMobility,
 * source/CALPHADMobility.h:31-160 and .cc:12-196, database block
 * MobilityParameters { Species<k> { Phase<L|A> { qA, qB, q0AB..q3AB } } }):
 *   Q_X(T)  = q[X][0] + R T ln q[X][1]  (+ qx[X][0] T ln T + qx[X][1] / T
 *             when has_extra), X = A (index 0), B (1);
 *   QQ_n(T) = qq[n][0] + R T ln qq[n][1]  (absent terms: (0, 1));
 *   dG = c0 Q_A + c1 Q_B + c0 c1 (QQ_0 + dc (QQ_1 + dc (QQ_2 + dc QQ_3))),
 *   dc = c0 - c1, c1 = 1 - c0;  mobility = exp(dG / RT) / RT. */
typedef struct ampe_calphad_mobility {
   double q[2][2];
   double qx[2][2];
   int has_extra;
   double qq[4][2];
} ampe_calphad_mobility;

typedef struct ampe_model_params {
   int ndim;              /* 2 or 3 */
   double dx[3];          /* mesh spacing */
   double temperature;    /* uniform temperature [K] */

   /* phase field */
   double epsilon_phi;    /* gradient energy coefficient epsilon_phi */
   double phi_well_scale; /* w: double-well height */
   double phi_mobility;   /* M_phi (uniform, SimpleQuatMobilityStrategy) */
   char phi_well_type;    /* 'd' double well, 's' single well */
   char energy_interp;    /* h(phi): 'p' pbg, 'h' harmonic, 'l' linear, ... */
   char conc_interp;      /* h(phi) used for c = (1-h) cL + h cS */
   char avg_type;         /* face average: 'a' arithmetic, 'h' harmonic */
   int flux_type;         /* enum ampe_flux_type */
   double aniso_nu;       /* 2-D: nu (epsilon_anisotropy); 3-D: eps4 */
   int aniso_knumber;     /* 2-D anisotropy fold k */

   /* orientation (quaternion) field */
   int evolve_quat;       /* 0/1 */
   int qlen;              /* quaternion components (4) */
   double H_parameter;    /* misorientation factor is 2*H */
   double epsilon_q;
   double quat_mobility;  /* M_q max */
   double min_quat_mobility;
   double quat_mobility_scale; /* alt_scale_factor of quatmobility */
   char quat_mobility_type;    /* 'p' pbg, 'e' exp, 'i' inverse */
   char orient_interp1;   /* p1(phi) of the 2H T p1 |grad q| term */
   char orient_interp2;   /* p2(phi) of the eps_q^2/2 p2 |grad q|^2 term */
   char floor_type;       /* 'm' max, 't' tanh, 's' sqrt */
   double gradient_floor;

   /* composition */
   int with_conc;         /* 0/1 */
   double D_liquid, Q0_liquid, D_solid, Q0_solid; /* D exp(-Q/RT) */
   double conc_mobility;  /* M_c multiplying div J */

   /* free energy / KKS */
   int fe_model;          /* enum ampe_free_energy */
   double vm_L, vm_A;     /* molar volumes [m^3/mol] (energy factor 1e-6/vm) */
   double quad_A[2];      /* quadratic well curvature, [0]=liquid [1]=solid */
   double quad_ceq[2];    /* quadratic equilibrium composition */
   double quad_m[2];      /* d ceq / dT */
   double quad_Tref;
   ampe_calphad_phase calphad[2]; /* [0]=liquid [1]=solid A */
   int newton_max_its;    /* KKS Newton iteration cap */
   double newton_tol;     /* KKS Newton residual tolerance */
   /* dilute binary solution (AMPE_FE_KKS_DILUTE, ConcentrationModel keys
    * keq, liquidus_slope, meltingT of tests/KKSdiluteBinaryKKS/test.input) */
   double dilute_keq;     /* partition coefficient k_e */
   double dilute_m;       /* liquidus slope [K per mole fraction], < 0 */
   double dilute_Tm;      /* melting temperature of the solvent [K] */
   /* phase mobility (QuatMobilityStrategy::computePhaseMobility) */
   int phi_mobility_type;         /* enum ampe_phi_mobility */
   double kim_interface_mobility; /* Kim: <= 0 infinite interface mobility */
   int kim_antitrapping;          /* Kim, finite mobility: anti-trapping form */
   /* composition right-hand side (ConcentrationModel rhs_form) */
   int conc_rhs_form;     /* enum ampe_conc_rhs_form */
   char diff_interp;      /* EBS phase weights h(phi) of the diffusivities
                           * (diffusion_interp_func_type, 'l' linear, 'p' pbg) */
   int ebs_stencil;       /* enum ampe_ebs_stencil (ConcentrationModel ebs_stencil) */
   /* multi-order phase field with three phases and a ternary quadratic model
    * (norderp, norderp_A of ModelParameters; ConcentrationModel { nspecies = 3,
    * model = "quadratic", Quadratic { A_liquid, Ceq_liquid, A_solid_A, ... } },
    * source/QuadraticFreeEnergyStrategyMultiOrderTernaryThreePhase.cc:24-80):
    * f_phase(c) = sum_k A[k] (c_k - ceq[k])^2 [J/mol] for phases L, A, B. */
   int norderp;           /* phase order parameters (> 1: multi-order; the last is the liquid) */
   int norderp_A;         /* orders 0 .. norderp_A-1 are grains of solid A, norderp_A ..
                           * norderp-2 grains of solid B */
   double vm_B;           /* molar volume of solid B [m^3/mol] */
   double tern_A[3][2];   /* curvatures per phase (L, A, B) and species */
   double tern_ceq[3][2]; /* equilibrium compositions per phase and species */
   double D_solid_B, Q0_solid_B; /* D exp(-Q/RT) of solid B (D_solid, Q0_solid: solid A) */
   /* EBS diffusivities (ConcentrationModel diffusion_type): enum
    * ampe_conc_diffusion.  Composition dependent: the atomic mobilities of the
    * two species in the liquid and in solid A, [phase][species]. */
   int conc_diffusion_type;
   ampe_calphad_mobility mobility[2][2];
} ampe_model_params;

/* EBS face diffusivities: temperature dependent (TbasedCompositionDiffusion-
 * Strategy: D exp(-Q/RT) per phase) or composition dependent
 * (MobilityCompositionDiffusionStrategy, the reference's default for the EBS
 * form: per phase the CALPHAD mobility times d2f/dc2 at the face-averaged
 * phase concentration, source/QuatModelParameters.cc:238-250). */
enum ampe_conc_diffusion {
   AMPE_CONC_DIFF_TEMPERATURE = 0,
   AMPE_CONC_DIFF_COMPOSITION = 1
};

/* Phase mobility: the uniform M_phi (SimpleQuatMobilityStrategy,
 * source/SimpleQuatMobilityStrategy.cc:46) or Kim's KKS mobility from the
 * liquid's d2f/dc2 (Kim, Acta Mater. 55 (2007) 4391; KimMobilityStrategy
 * InfMob / FiniteMob / FiniteMobAntiTrap, source/KimMobilityStrategy*.cc,
 * factor of source/KKStools.cc:13-30), evaluated per cell from (cL, cS). */
enum ampe_phi_mobility { AMPE_PHI_MOB_CONSTANT = 0, AMPE_PHI_MOB_KIM = 1 };
/* Composition RHS: KKS (KKSCompositionRHSStrategy, flux from D_c grad c and
 * D_phi grad phi) or EBS (EBSCompositionRHSStrategy, flux sum over phases of
 * D_phase grad c_phase, source/EBSCompositionRHSStrategy.cc:183-328). */
enum ampe_conc_rhs_form { AMPE_CONC_RHS_KKS = 0, AMPE_CONC_RHS_EBS = 1 };
/* EBS flux stencil (EBSCompositionRHSStrategy, source/EBSCompositionRHSStrategy.cc
 * :149-154; QuatModelParameters ebs_stencil): "regular" add_flux
 * (2d/flux.m4:11-66), "isotropic" add_flux_iso (2d/flux.m4:69-135, 3d/flux.m4
 * :103-232; transverse neighbours weighted 1/12, the face itself 10/12 in 2-D,
 * 8/12 in 3-D), "order4" add_flux_4th (2d/flux.m4:139-242, 3d/flux.m4:235-420;
 * fourth-order face gradients, second order on physical-boundary faces; the
 * integrator's level then has two ghost cells, as nghosts_required gives,
 * source/QuatModelParameters.h:305-311). */
enum ampe_ebs_stencil { AMPE_EBS_REGULAR = 0, AMPE_EBS_ISOTROPIC = 1, AMPE_EBS_ORDER4 = 2 };
enum ampe_precision { AMPE_FP64 = 0, AMPE_FAC_FP32 = 1 };
enum ampe_kks_guess { AMPE_GUESS_PREVIOUS = 0, AMPE_GUESS_SURROGATE = 1 };
/* Block solver inside each FAC cycle: the reference's single-level red-black
 * coarse solve (EllipticFACOps "redblack", source/EllipticFACOps.cc:1366-1370)
 * or a geometric multigrid V-cycle below the level (the role of hypre PFMG,
 * the reference's default coarse solver). */
enum ampe_precond { AMPE_PRECOND_REDBLACK = 0, AMPE_PRECOND_MG = 1 };
/* SPGMR orthogonalisation (SUNLinSol_SPGMR gstype): modified Gram-Schmidt
 * (the reference's choice, source/QuatIntegrator.cc:1571-1596) or classical
 * Gram-Schmidt with one conditional re-orthogonalisation (SUNClassicalGS). */
enum ampe_gstype { AMPE_GS_MODIFIED = 0, AMPE_GS_CLASSICAL = 1 };

typedef struct ampe_solver_params {
   double dt;             /* fixed BDF step (initial = max step) */
   int max_order;         /* 1 or 2 */
   double rtol, atol;     /* CVODE scalar tolerances */
   int maxl;              /* SPGMR Krylov dimension (max_krylov_dimension) */
   int max_newton;        /* Newton iterations per attempt (CVODE maxcor) */
   double nlscoef;        /* Newton convergence coefficient */
   double eplifac;        /* linear / nonlinear tolerance ratio */
   int msbp;              /* steps between preconditioner setups */
   int lag_quat_sidegrad; /* lag |grad q| face norms in Jv evaluations */
   int fac_max_cycles;    /* phase & composition FAC cycles */
   int fac_sweeps;        /* RBGS sweeps per cycle (coarse_solver_max_iterations) */
   double fac_tol;        /* RBGS early exit (coarse_solver_tolerance, <0: off) */
   int quat_max_cycles;   /* orientation block cycles */
   int quat_sweeps;       /* orientation block RBGS sweeps per cycle */
   int precision;         /* enum ampe_precision */
   int kks_guess;         /* enum ampe_kks_guess */
   int precond;           /* enum ampe_precond */
   int mg_max_levels;     /* multigrid levels including the fine one (<= 0: all) */
   int mg_coarse_sweeps;  /* red-black sweeps on the coarsest multigrid level */
   int gstype;            /* enum ampe_gstype (classical: maxl <= 14) */
   int graph_psolve;      /* 1: the FP32 preconditioner solve of the Krylov loop (the
                           * multigrid V-cycles or red-black sweeps of every block) is
                           * captured once per gamma into a HIP graph and replayed with
                           * one launch per solve -- the role of a captured V-cycle; needs
                           * precision = FP32, gstype = classical and one cycle per
                           * block (fac_max_cycles = quat_max_cycles = 1), else ignored */
   double dt_init;        /* first step (0 < dt_init < dt; else dt); the step doubles
                           * after every accepted step up to dt, as after a cut --
                           * the small automatic first step CVODE takes when the
                           * initial step size is 0 (QuatIntegrator's default
                           * d_previous_timestep, source/QuatIntegrator.cc:135,1575) */
   int adaptive;          /* 0: the imposed step above (BDF1, then variable-step BDF2).
                           * 1: CVODE's adaptive BDF as QuatIntegrator configures it
                           * (source/QuatIntegrator.cc:1571-1596): fixed-leading-
                           * coefficient BDF of order 1..max_order (<= 2) on a
                           * Nordsieck history, local error test on the projected
                           * correction (dsm = ||acor||_wrms tq2 <= 1), step and
                           * order selection after each step, first step dt_init
                           * (else dt), largest step dt (max_step_size) */
   int mg_quat_natural;   /* FP32 multigrid: 0 (default) stores the orientation block's
                           * multigrid vectors row parity-split (inside each ghosted row
                           * the even columns, then the odd ones), so one red-black
                           * colour reads and writes whole lines; 1 keeps the natural
                           * layout.  Same values either way. */
} ampe_solver_params;

/* Per-run counters (mirror QuatIntegrator.cc:2161-2170 nni/nli/nfe/npe/nps). */
typedef struct ampe_stats {
   long long nst, nni, nli, nfe, npe, nps, ncfn, nlcf;
   long long kks_calls, kks_cells, kks_iters, kks_fallbacks;
   long long kks_surrogate_starts; /* cells whose Newton started from the surrogate */
   double last_acnrm; /* WRMS norm of the last step's Newton correction */
   double t;          /* simulated time reached */
   double h;          /* next step size (dt unless a failure cut it) */
   long long netf;    /* local error test failures (adaptive mode) */
   long long q;       /* order of the last accepted step */
} ampe_stats;

/* ---------------------------------------------------------------- runtime */
const char* ampe_last_error(void);
int ampe_version(void);

/* --------------------------------------------------------------- levels */
/* A patch level: uneven boxes, one ghost width, optional periodic domain.
 * Replaces the SAMRAI PatchLevel + ghost RefineSchedule role used by
 * QuatIntegrator::fillScratch (source/QuatIntegrator.cc:2873-2957). */
typedef struct ampe_level ampe_level;
int ampe_level_create(int ndim, const ampe_box* boxes, int npatch, int ghost,
                      const ampe_box* periodic_domain, ampe_level** out);
int ampe_level_destroy(ampe_level* level);
/* ncell_ghost: cells of one component incl. ghosts; ncell: interior cells;
 * nside[d]: faces of direction d (all patches). */
int ampe_level_sizes(const ampe_level* level, long long* ncell_ghost,
                     long long* ncell, long long nside[3]);
/* Host copy of per-patch offsets: cell_off[npatch], side_off[3*npatch]. */
int ampe_level_patch_offsets(const ampe_level* level, long long* cell_off,
                             long long* side_off);
/* Strides of each patch's ghosted cell block: cell (i, j, k), interior
 * indices from 0, ghost width g (gz = g in 3-D, 0 in 2-D), sits at
 * cell_off + (i + g) + (j + g) * sy + (k + gz) * sz.  sy >= nx + 2g (rows may
 * be padded); sy[npatch], sz[npatch]. */
int ampe_level_patch_strides(const ampe_level* level, long long* sy, long long* sz);
/* Fill ghost cells of ncomp component blocks from neighbours / periodic
 * images (SAMRAI RefineSchedule::fillData on one level). */
int ampe_ghost_fill(const ampe_level* level, double* field, int ncomp,
                    ampe_stream_t stream);

/* ------------------------------------------------------------ RHS kernels */
/* |grad q| at cells (quat_grad_modulus_type "cells") and, if face_norm2 is
 * not NULL, |grad q|^2 on every face from the full side gradient.
 * Replaces QuatGradStrategy::computeDiffs/computeGradCell/computeGradSide/
 * computeGradModulus (source/QuatGradStrategy.h:32-73; F77 quatdiffs
 * source/fortran/2d/quatdiffs.m4:56, quatgrad_cell quatgrad.m4:12,
 * quatgrad_side quatgrad.m4:157, quatgrad_modulus quatgrad.m4:882;
 * face norm as in compute_face_coef2d quatfacops.m4:83-92). q needs filled
 * ghosts; gmod is a cell field (interior written). */
int ampe_quat_grad(const ampe_level* level, const ampe_model_params* p,
                   const double* q, double* gmod, double* face_norm2,
                   ampe_stream_t stream);

/* Phase RHS: flux (Simple / Isotropic / Anisotropy), div, well, orientation
 * terms, KKS driving force, times M_phi -- fused.  Replaces
 * PhaseRHSStrategyWithQ::evaluateRHS (source/PhaseRHSStrategyWithQ.cc:90-312)
 * with PhaseFluxStrategy::computeFluxes (source/PhaseFluxStrategy.h:26) and
 * FreeEnergyStrategy::addDrivingForce (source/FreeEnergyStrategy.h:44-60).
 * phi, q need filled ghosts; cl, cs are cell fields; rhs: cell field. */
int ampe_phase_rhs(const ampe_level* level, const ampe_model_params* p,
                   const double* phi, const double* q, const double* cl,
                   const double* cs, double* rhs, ampe_stream_t stream);

/* Orientation RHS with projection: face coefficients from lagged face
 * |grad q|^2, flux fc * dq/dx, divergence, lambda, -M_q(...).  Replaces
 * QuatSysSolver::evaluateRHS -> QuatFACOps::evaluateRHS
 * (source/QuatFACOps.cc:1861-1890, 1958-2136; F77 compute_face_coef2d,
 * compute_flux2d_from_gradq, compute_lambda_flux2d, add_quat_proj_op2d). */
int ampe_quat_rhs(const ampe_level* level, const ampe_model_params* p,
                  const double* phi, const double* q,
                  const double* face_norm2, double* rhs, ampe_stream_t stream);

/* The same in two passes, as QuatFACOps does it: the face coefficients of
 * every face from the face |grad q|^2 (QuatFACOps::computeFaceCoefs,
 * source/QuatFACOps.cc:735-829, F77 compute_face_coef2d/3d), then the RHS
 * from them (evaluateRHS with the coefficients, source/QuatFACOps.cc
 * :1861-1890).  Same values as ampe_quat_rhs; fc is a side field. */
int ampe_quat_face_coefs(const ampe_level* level, const ampe_model_params* p,
                         const double* phi, const double* n2, double* fc,
                         ampe_stream_t stream);
int ampe_quat_rhs_fc(const ampe_level* level, const ampe_model_params* p,
                     const double* phi, const double* q, const double* fc, double* rhs,
                     ampe_stream_t stream);

/* CVODE projection of q onto |q| = 1 (QuatIntegrator::applyProjection,
 * source/QuatIntegrator.cc:3911-3965; F77 project2d quatfacops.m4:940). */
int ampe_quat_project(const ampe_level* level, const ampe_model_params* p,
                      const double* q, double* corr, double* err,
                      ampe_stream_t stream);

/* Composition face coefficients D_c (pfm diffusion) and D_phi (phase
 * coupling).  Replaces KKSCompositionRHSStrategy::setDiffusionCoeff
 * (source/KKSCompositionRHSStrategy.cc:72-373; F77
 * concentration_pfmdiffusion concentrationdiffusion.m4:14). */
int ampe_conc_coeffs(const ampe_level* level, const ampe_model_params* p,
                     const double* phi, const double* cl, const double* cs,
                     double* Dc, double* Dphi, ampe_stream_t stream);

/* Composition RHS: J = (Dc dc + Dphi dphi)/dx, rhs = M_c div J.  Replaces
 * KKSCompositionRHSStrategy::computeFluxOnPatch + computerhsconcentration
 * (source/KKSCompositionRHSStrategy.cc:377-441, QuatIntegrator.cc:2578-2690). */
int ampe_conc_rhs(const ampe_level* level, const ampe_model_params* p,
                  const double* c, const double* phi, const double* Dc,
                  const double* Dphi, double* rhs, ampe_stream_t stream);

/* Per-cell KKS phase concentrations over the ghost box, Newton from
 * (cl_ref, cs_ref) (or the closed form for quadratic wells).  Replaces
 * PhaseConcentrationsStrategy::computePhaseConcentrations
 * (source/PhaseConcentrationsStrategy.cc:26-127) with the CALPHAD binary and
 * quadratic OnPatch strategies.  *max_iters (host) receives the largest
 * Newton iteration count, or -1 if a cell did not converge. */
int ampe_kks_solve(const ampe_level* level, const ampe_model_params* p,
                   const double* phi, const double* c, const double* cl_ref,
                   const double* cs_ref, double* cl, double* cs,
                   int* max_iters, ampe_stream_t stream);

/* EBS composition model, two phases: face diffusivities D_L = (1-h) D_L0
 * exp(-Q_L/RT), D_A = h D_S0 exp(-Q_S/RT), h = diff_interp(avg(phi)) (the
 * temperature-dependent strategy's concentration_pfmdiffusion_of_temperature,
 * source/fortran/2d/concentrationdiffusion.m4:341-428 via
 * TbasedCompositionDiffusionStrategy::setDiffusion), then
 * rhs = M_c div(D_L grad cL + D_A grad cS) (EBSCompositionRHSStrategy::
 * computeFluxOnPatch with add_flux, source/fortran/2d/flux.m4:11-66, and
 * computerhsconcentration).  cl, cs need filled ghosts (the KKS solve
 * covers them); DL, DA are side fields. */
/* EBS face diffusivities of the composition-dependent model
 * (MobilityCompositionDiffusionStrategy::setDiffusion,
 * source/MobilityCompositionDiffusionStrategy.cc:97-150, 194-329 and
 * setPFMDiffOnPatch :331-611; computeDiffusionMobilityBinaryPhase,
 * source/CALPHADMobility.cc:200-218): on each face the phase concentrations
 * are averaged arithmetically, c = (c_i + c_i-1) / 2, the phase's diffusivity
 * is M(c) d2f/dc2(c) with M = c (1 - c) (c m_1 + (1 - c) m_0) 1e12 um^2 from
 * the two species' atomic mobilities m_k (ampe_calphad_mobility, model
 * p->mobility) and d2f/dc2 in J/mol, and D_L = (1 - h) D_L(c_L),
 * D_A = h D_A(c_A), h = diff_interp(avg(phi)).  cl, cs with filled ghosts. */
int ampe_conc_coeffs_ebs_mobility(const ampe_level* level, const ampe_model_params* p,
                                  const double* phi, const double* cl, const double* cs,
                                  double* DL, double* DA, ampe_stream_t stream);

int ampe_conc_coeffs_ebs(const ampe_level* level, const ampe_model_params* p,
                         const double* phi, double* DL, double* DA,
                         ampe_stream_t stream);
int ampe_conc_rhs_ebs(const ampe_level* level, const ampe_model_params* p,
                      const double* cl, const double* cs, const double* DL,
                      const double* DA, double* rhs, ampe_stream_t stream);
/* Multi-order phase field (norderp > 1, two phases): MultiOrderRHSStrategy::
 * evaluateRHS (source/MultiOrderRHSStrategy.cc:146-259) with the simple
 * gradient flux of epsilon_phi / sqrt(2) (source/PhaseFluxStrategyFactory.h
 * :19-22) and computerhsmultiorder (source/fortran/2d/multiorder.m4:16-105,
 * 3d/multiorder.m4): for each order p,
 *   rhs_p = M [div(eps^2 grad phi_p) - m (phi_p^3 - phi_p
 *           + 2 gamma phi_p (sum_q phi_q^2 - phi_p))],
 * m = 8 phi_well_scale, gamma = 1.5 (the reference subtracts phi_p, not
 * phi_p^2, inside the sum).  phi and rhs hold norder components of the
 * level's ghosted size (phi's ghosts filled); mobility is a cell field or
 * NULL for the uniform phi_mobility.  The free-energy driving force of
 * three-phase models is not included.  Needs flux_type = AMPE_FLUX_SIMPLE. */
int ampe_phase_rhs_multiorder(const ampe_level* level, const ampe_model_params* p,
                              int norder, const double* phi, const double* mobility,
                              double* rhs, ampe_stream_t stream);
/* Three phases (liquid = the last order, solid A = orders 0..norderp_A-1,
 * solid B = the rest) with the ternary quadratic model
 * (fe_model = AMPE_FE_QUADRATIC_TERNARY_3PH, p->norderp >= 3 orders, two
 * species); phase weights h = (phi_L^2, sum_A phi^2, sum_B phi^2) / sum phi^2.
 * ph: six ghosted fields cL0, cL1, cA0, cA1, cB0, cB1.
 * ampe_kks_ternary3: QuadraticEquilibriumThreePhasesTernaryStrategyMultiOrder::
 * computePhaseConcentrationsOnPatch (source/QuadraticEquilibriumThreePhases
 * TernaryStrategyMultiOrder.cc:49-181) over the ghost box: for
 * f_a(c) = sum_k A_ak (c_k - e_ak)^2 the KKS conditions (one potential
 * mu_k = 2 A_ak (x_ak - e_ak), c_k = sum_a g(h_a) x_ak, g = conc_interp) in
 * closed form; c holds two fields.
 * ampe_phase_rhs_ternary3: ampe_phase_rhs_multiorder's terms minus Moelans'
 * driving force 2 phi_p [(1 - h_a) w_a - sum_(b != a) h_b w_b] / sum phi^2,
 * w = f - mu . c (QuadraticFreeEnergyStrategyMultiOrderTernaryThreePhase::
 * addDrivingForceOnPatch, source/QuadraticFreeEnergyStrategyMultiOrder
 * TernaryThreePhase.cc:327-520), then times the mobility.
 * ampe_ebs3_coeffs: face diffusivities D_a = g_diff(sum over a's orders of
 * the face average) D0_a exp(-Q_a/RT) (concentration_pfmdiffusion_of_
 * temperature_threephases, source/fortran/2d/concentrationdiffusion.m4
 * :523-640; D_solid/Q0_solid for A, D_solid_B/Q0_solid_B for B).
 * ampe_conc_rhs_ebs3: rhs_k = M_c div(D_L grad cL_k + D_A grad cA_k +
 * D_B grad cB_k) for both species (EBSCompositionRHSStrategy::
 * computeFluxOnPatch with add_flux, source/EBSCompositionRHSStrategy.cc
 * :183-328); rhs holds two fields. */
int ampe_kks_ternary3(const ampe_level* level, const ampe_model_params* p, const double* phi,
                      const double* c, double* ph, ampe_stream_t stream);
int ampe_phase_rhs_ternary3(const ampe_level* level, const ampe_model_params* p,
                            const double* phi, const double* ph, const double* mobility,
                            double* rhs, ampe_stream_t stream);
int ampe_ebs3_coeffs(const ampe_level* level, const ampe_model_params* p, const double* phi,
                     double* DL, double* DA, double* DB, ampe_stream_t stream);
int ampe_conc_rhs_ebs3(const ampe_level* level, const ampe_model_params* p, const double* ph,
                       const double* DL, const double* DA, const double* DB, double* rhs,
                       ampe_stream_t stream);
/* Projection onto sum_p phi_p = 1 (projectphi, source/fortran/2d/phase.m4
 * :51-110, via MultiOrderRHSStrategy::projectPhases, source/MultiOrderRHSStrategy.cc
 * :262-300): corr_p = (phi_p - (sum phi - 1)/norder) - phi_p, and err loses
 * its mean over the orders, per interior cell. */
int ampe_project_phases(const ampe_level* level, int norder, const double* phi,
                        double* corr, double* err, ampe_stream_t stream);
/* The same with the fourth-order stencil (add_flux_4th, 2d/flux.m4:139-242,
 * 3d/flux.m4:235-420): face flux D (15 (c_h - c_m) - (c_hh - c_mm)) / (12 dx),
 * but D (c_h - c_m) / dx on faces of the physical boundary (the sides of the
 * level's bounding box of a non-periodic level: computeFluxOnPatch's physbc,
 * source/EBSCompositionRHSStrategy.cc:236-255).  The level needs ghost >= 2
 * and cl, cs two filled ghost layers. */
int ampe_conc_rhs_ebs_4th(const ampe_level* level, const ampe_model_params* p,
                          const double* cl, const double* cs, const double* DL,
                          const double* DA, double* rhs, ampe_stream_t stream);
/* The same with the isotropic stencil (add_flux_iso): the face diffusivities,
 * D_L = (1 - h(phi_f)) D_L0 e^{-Q_L/RT}, D_A = h(phi_f) D_S0 e^{-Q_S/RT} with
 * phi_f the face average of phi_d, are evaluated for every face the stencil
 * touches, including those beside the patch (phi_d needs filled ghosts). */
int ampe_conc_rhs_ebs_iso(const ampe_level* level, const ampe_model_params* p,
                          const double* phi_d, const double* cl, const double* cs,
                          double* rhs, ampe_stream_t stream);

/* Event timing of one kernel class inside whatever runs while it is enabled
 * (used by the benchmark to measure a kernel live inside the timed steps).
 * kernel: enum ampe_timed_kernel.  ampe_kernel_timer_read waits for the
 * recorded launches, returns their total duration [ms], count and algorithmic
 * bytes, and clears them.  No reference counterpart (the reference times whole
 * phases with SAMRAI timers); process-wide, not thread-safe. */
enum ampe_timed_kernel { AMPE_TIMED_QUAT_RBGS = 0 };
int ampe_kernel_timer_enable(int kernel, int on);
int ampe_kernel_timer_read(int kernel, double* ms, long long* launches, double* bytes);

/* Grain identification (Grains::findAndNumberGrains, source/Grains.cc
 * :263-645): cells with phase >= threshold, 4-connected (2-D) or 6-connected
 * (3-D) across patch and periodic boundaries, numbered 0, 1, ... in the order
 * of each grain's smallest index in the level's bounding box (x fastest);
 * labels: int, interior cells per patch as in ampe_tag_cells, -1 outside
 * grains.  Synchronises the stream; *ngrains = number of grains. */
int ampe_grains_label(const ampe_level* level, const double* phase, double threshold,
                      int* labels, int* ngrains, ampe_stream_t stream);

/* AMR gradient detector: tags = 1 where |grad phi * dx|^2 >= tol^2, else 0
 * (QuatModel::tagGradientDetectorCells source/QuatModel.cc:3476-3602,
 * F77 tagfromgrads quattag.m4:9).  tags: int, interior cells per patch. */
int ampe_tag_cells(const ampe_level* level, const ampe_model_params* p,
                   const double* phi, double tol, int* tags,
                   ampe_stream_t stream);

/* Box generation from tags: the role of SAMRAI's BergerRigoutsos in
 * GriddingAlgorithm::regridAllFinerLevels (source/PFModel.cc:258,565-590),
 * with the tag buffering of bufferTagsOnLevel (source/PFModel.cc:336).  The
 * clustering is Berger & Rigoutsos (1991) as restated in
 * ampe_amd/csrc/cluster.hip (SAMRAI's own variant is not reproduced); boxes
 * are disjoint, cover every buffered tag and lie in the level's bounding box. */
typedef struct ampe_cluster_params {
   double efficiency_tol; /* accept a box when tagged / cells >= this
                           * (BergerRigoutsos efficiency_tolerance) */
   int min_size[3];       /* smallest extent of a cut piece when a box is split; accepted
                             boxes are the tags' bounding boxes and may be smaller (SAMRAI's
                             smallest_patch_size, which grows boxes, is not reproduced) */
   int max_size[3];       /* accepted boxes are chopped to at most this extent
                           * (largest_patch_size; <= 0: no limit) */
   int tag_buffer;        /* tags grown by this many cells in every direction */
} ampe_cluster_params;
/* Host mask (x fastest over the domain box, nonzero = tagged). */
int ampe_cluster_mask(int ndim, const ampe_box* domain, const unsigned char* mask,
                      const ampe_cluster_params* cp, ampe_box* boxes, int max_boxes,
                      int* nboxes);
/* Device tags of a level (as ampe_tag_cells writes them); synchronises the
 * stream. */
int ampe_cluster_tags(const ampe_level* level, const int* tags, const ampe_cluster_params* cp,
                      ampe_box* boxes, int max_boxes, int* nboxes, ampe_stream_t stream);

/* -------------------------------------------------------- FAC operators */
/* Elliptic operator A u = M div(D grad u) + C u on one level
 * (EllipticFACOps, source/EllipticFACOps.h:103).  Each coefficient is a
 * field or, when the pointer is NULL, the constant. */
typedef struct ampe_elliptic_op {
   const double* D;  /* side field or NULL */
   double D_const;
   const double* C;  /* cell field (interior) or NULL */
   double C_const;
   const double* M;  /* cell field (interior) or NULL */
   double M_const;
   double dx[3];     /* mesh spacing */
} ampe_elliptic_op;

/* PhaseFACOps::setOperatorCoefficients / setC (source/PhaseFACOps.cc:33-183):
 * C = 1 + gamma*M*w*g''(phi) written to C (cell field); M is the phase
 * mobility field, or NULL for the uniform M_phi. */
int ampe_fac_phase_coeffs(const ampe_level* level, const ampe_model_params* p,
                          double gamma, const double* phi, const double* M,
                          double* C, ampe_stream_t stream);

/* Phase mobility field over interior cells from the phase concentrations
 * (QuatMobilityStrategy::computePhaseMobility as called by
 * QuatIntegrator::computeMobilities, source/QuatIntegrator.cc:2959-2992; Kim
 * forms of source/KimMobilityStrategy*.cc).  cl, cs may be NULL for the
 * uniform mobility. */
int ampe_phase_mobility(const ampe_level* level, const ampe_model_params* p,
                        const double* cl, const double* cs, double* M,
                        ampe_stream_t stream);

/* num_sweeps red-black Gauss-Seidel sweeps on A u = rhs (u ghosted, ghost
 * fill before each colour), with early exit when max|res| <= tol (tol >= 0).
 * EllipticFACOps::smoothErrorByRedBlack (source/EllipticFACOps.cc:1067-1230)
 * with efo_rbgswithfluxmax* (source/fortran/2d/ellipticfacops.m4:60-345).
 * *maxres (host, may be NULL) receives the last sweep's max |residual|. */
int ampe_fac_smooth(const ampe_level* level, const ampe_elliptic_op* op,
                    const double* rhs, double* u, int num_sweeps, double tol,
                    double* maxres, ampe_stream_t stream);

/* res = rhs - A u (efo_compres*sca, source/fortran/2d/ellipticfacops.m4:346)
 * after a ghost fill of u; *rms (host, may be NULL) = RMS norm of res. */
/* num_sweeps red-black sweeps of the orientation block
 * (I + gamma sqrt(M) L sqrt(M)) z = rhs on ncomp (<= 4) component blocks:
 * fc face coefficients (side field), sm = sqrt(M) (cell field, ghosts
 * filled), dx the level spacing; z ghosted, updated in place.  fused = 0: a
 * red then a black launch per sweep with ghost fills between; fused = 1: the
 * one-pass sweep (same values).  The role of QuatFACOps::smoothError /
 * QuatLevelSolver (source/QuatFACOps.cc:1009-1091), which the reference
 * gives to hypre; this build smooths the same symmetrised matrix. */
int ampe_quat_smooth(const ampe_level* level, const double* fc, const double* sm,
                     double gamma, const double dx[3], int ncomp, const double* rhs,
                     double* z, int num_sweeps, int fused, ampe_stream_t stream);

int ampe_fac_residual(const ampe_level* level, const ampe_elliptic_op* op,
                      const double* rhs, double* u, double* res, double* rms,
                      ampe_stream_t stream);

/* Coarse right-hand side of a FAC block's multigrid: b (on coarse, the fine
 * boxes halved) = the mean over each coarse cell's children of the fine
 * residual rhs - A u (u ghost-filled here): computeCompositeResidualOnLevel
 * followed by restrictResidual (source/EllipticFACOps.cc:1593, 2145-2276,
 * 948-1041).  red_only = 0: every cell's residual to res, then the
 * restriction.  red_only = 1: u comes out of red-black sweeps ending with the
 * black colour, whose residual vanishes on black cells, so only the red
 * cells' residuals are formed, per coarse cell (res unused); needs a periodic
 * fine level with even, non-negative patch corners. */
int ampe_fac_residual_restrict(const ampe_level* coarse, const ampe_level* fine,
                               const ampe_elliptic_op* op, const double* rhs, double* u,
                               double* res, double* b, int red_only, ampe_stream_t stream);

/* ------------------------------------------------ level transfer (ratio 2) */
/* coarse must hold the boxes of fine halved, patch by patch (even lower
 * corners and extents).  Cell fields of ncomp components; side fields of one.
 * Replace EllipticFACOps::restrictSolution / restrictResidual /
 * prolongErrorAndCorrect (source/EllipticFACOps.cc:948-1041): SAMRAI
 * CONSERVATIVE_COARSEN on cells (mean of the 2^d children) and faces (mean of
 * the fine faces covering a coarse face), and refinement added to the fine
 * field: order 0 = CONSTANT_REFINE (the reference default), order 1 =
 * cell-centred tri/bilinear interpolation (needs filled coarse ghosts). */
int ampe_coarsen_cell(const ampe_level* coarse, const ampe_level* fine, int ncomp,
                      const double* f, double* c, ampe_stream_t stream);
int ampe_coarsen_side(const ampe_level* coarse, const ampe_level* fine, const double* f,
                      double* c, ampe_stream_t stream);
int ampe_refine_add(const ampe_level* fine, const ampe_level* coarse, int ncomp, int order,
                    const double* c, double* f, ampe_stream_t stream);

/* Coarse right-hand side of the orientation block's multigrid: b (on coarse,
 * the fine boxes halved) = the mean over each coarse cell's children of the
 * fine residual rhs - A z, A = I + gamma sqrt(M) L sqrt(M) as in
 * ampe_quat_smooth (z ghost-filled here).  The residual and restriction steps
 * of the orientation FAC cycle, QuatFACOps::computeCompositeResidualOnLevel
 * and restrictResidual (source/QuatFACOps.cc:1588-1754, 939-950).
 * red_only = 0: the residual of every cell goes to res (ncomp fine component
 * blocks), then the restriction.  red_only = 1: z comes out of red-black
 * sweeps ending with the black colour, whose residual vanishes on black
 * cells, so only the red cells' residuals are formed, per coarse cell, and
 * res is not used; needs a periodic fine level with even, non-negative patch
 * corners (one checkerboard) and ncomp 1, 2 or 4. */
int ampe_quat_residual_restrict(const ampe_level* coarse, const ampe_level* fine,
                                const double* fc, const double* sm, double gamma,
                                const double dx[3], int ncomp, const double* rhs, double* z,
                                double* res, double* b, int red_only, ampe_stream_t stream);
/* Orientation field of qlen <= 4 components: coarsened by QuatWeightedAverage
 * (source/QuatWeightedAverage.cc:100-171, F77 quatcoarsen: volume average of
 * the children, then q/|q|) and refined (set, not added) by QuatLinearRefine
 * (source/QuatLinearRefine.cc:107-174, F77 quatlinrefine: bi/trilinear
 * interpolation of the coarse cells, then q/|q|; coarse ghosts filled).  The
 * reference's symmetry rotation with index 0 is applied as the identity (its
 * qlen 1 / 2 behaviour; for qlen 4 the reference stops).  dx_fine: fine mesh
 * spacing.  These are the symmetry-aware operators (symmetry_aware = true);
 * with the reference's default (false, source/QuatModel.cc:505, 946-948) an
 * orientation field is refined by plain LINEAR_REFINE, ampe_linear_refine with
 * ncomp = qlen, which does not normalise -- the only qlen 4 case. */
int ampe_quat_coarsen(const ampe_level* coarse, const ampe_level* fine, const double* dx_fine,
                      int qlen, const double* qf, double* qc, ampe_stream_t stream);
int ampe_quat_refine(const ampe_level* fine, const ampe_level* coarse, const double* dx_fine,
                     int qlen, const double* qc, double* qf, ampe_stream_t stream);
/* Cell-centred linear refinement (set, not added) of ncomp scalar components:
 * the (bi/tri)linear interpolation of ampe_quat_refine without the
 * normalisation, the role of SAMRAI's "LINEAR_REFINE" that QuatModel registers
 * for the phase, composition and temperature (source/QuatModel.cc:924-955) when
 * a new fine level is filled from the coarse one (SAMRAI is absent: its
 * operator is restated from quatlinrefine's interpolation).  Coarse ghosts
 * filled; fine and coarse paired patch by patch as for ampe_quat_refine. */
int ampe_linear_refine(const ampe_level* fine, const ampe_level* coarse, const double* dx_fine,
                       int ncomp, const double* c, double* f, ampe_stream_t stream);

/* ----------------------------------------------------- vector reductions */
/* Dot product of two ncomp-component cell fields over interior cells
 * (SAMRAIVectorReal::dot as reached through Sundials_SAMRAIVector::dotWith,
 * source/samrai/Sundials_SAMRAIVector.cc:156, without the cell-volume factor,
 * which the caller applies).  Deterministic fixed-order reduction. */
int ampe_vec_dot(const ampe_level* level, int ncomp, const double* x,
                 const double* y, double* out, ampe_stream_t stream);
/* Weighted RMS norm sqrt(sum (x w)^2 / N) (Sundials_SAMRAIVector::
 * weightedRMSNorm, source/samrai/Sundials_SAMRAIVector.cc:177). */
int ampe_vec_wrms(const ampe_level* level, int ncomp, const double* x,
                  const double* w, double* out, ampe_stream_t stream);

/* ------------------------------------------- temperature-dependent CALPHAD */
/* One end-member Gibbs energy on nrange temperature intervals
 * [Tc[k], Tc[k+1]):  G(T) = a + b T + c T ln T + d2 T^2 + d3 T^3 + d7 T^7
 *                          + dm1 / T + dm9 T^-9                  [J/mol]
 * (the thermodynamic_data/calphadAuNi.dat:5-45 database form). */
typedef struct ampe_calphad_gfun {
   int nrange;
   double Tc[5];
   double coef[4][8]; /* per interval: a, b, c, d2, d3, d7, dm1, dm9 */
} ampe_calphad_gfun;

/* Binary CALPHAD phases with temperature dependence: end members g[phase]
 * [species] (phase 0 = liquid, 1 = solid A; species 0 = A, 1 = B) and
 * Redlich-Kister L_i(T) = L[phase][i][0] + L[phase][i][1] T
 * (calphadAuNi.dat:47-63). */
typedef struct ampe_calphad_tdep {
   ampe_calphad_gfun g[2][2];
   double L[2][4][2];
} ampe_calphad_tdep;

/* Evaluate a tdep table at temperature T into two ampe_calphad_phase records
 * (host helper, same arithmetic as the device path). */
int ampe_calphad_at(const ampe_calphad_tdep* tab, double T,
                    ampe_calphad_phase out[2]);

/* ------------------------------------------------------ KKS surrogate MLP */
/* No counterpart in the reference: a small MLP n_in -> H -> H -> n_out with
 * tanh hidden layers, BF16 weights and activations on the gfx950 matrix cores
 * with FP32 accumulation, that supplies the initial guess of the per-cell KKS
 * Newton solve in place of the guess the reference takes from the previous
 * step's phase concentrations
 * (source/CALPHADequilibriumPhaseConcentrationsStrategy.cc:368-369).
 * Parameters are FP32 masters, flattened as W1[H][n_in], b1[H], W2[H][H],
 * b2[H], W3[n_out][H], b3[n_out].  Network inputs and outputs are normalised:
 * x_n = (x - in_shift) * in_scale, y = y_n * out_scale + out_shift. */
typedef struct ampe_surrogate ampe_surrogate;
/* n_in <= 4, n_hidden in {64, 128}, n_out <= 4; Glorot-uniform weights from
 * seed, zero biases, identity scaling. */
int ampe_surrogate_create(int n_in, int n_hidden, int n_out,
                          unsigned long long seed, ampe_surrogate** out);
int ampe_surrogate_destroy(ampe_surrogate* s);
int ampe_surrogate_nparams(const ampe_surrogate* s, long long* n);
/* Host FP32 parameter vectors of length nparams. */
int ampe_surrogate_set_params(ampe_surrogate* s, const float* params);
int ampe_surrogate_get_params(const ampe_surrogate* s, float* params);
int ampe_surrogate_set_scaling(ampe_surrogate* s, const double* in_shift,
                               const double* in_scale, const double* out_shift,
                               const double* out_scale);
/* Forward pass on normalised inputs: x device float[n_in][n] (SoA), y device
 * float[n_out][n]. */
int ampe_surrogate_forward(const ampe_surrogate* s, long long n,
                           const float* x, float* y, ampe_stream_t stream);
/* Mean-squared-error loss L = (1/n) sum_samples sum_outputs (y - t)^2 on
 * normalised data and its gradient (device float[nparams]); *loss on host. */
int ampe_surrogate_loss_grad(ampe_surrogate* s, long long n, const float* x,
                             const float* t, float* grad, double* loss,
                             ampe_stream_t stream);

typedef struct ampe_train_params {
   long long batch;   /* samples per Adam step */
   int steps;         /* Adam steps */
   double lr0, lr1;   /* learning rate, decayed geometrically lr0 -> lr1 */
   double beta1, beta2, eps;
} ampe_train_params;
/* Adam training on raw (unnormalised) FP64 device data: x double[n_in][n],
 * t double[n_out][n]; the object's scaling normalises both.  Minibatches
 * walk the data in order (draw it in random order).  loss_hist (host, may
 * be NULL) receives the minibatch loss of every step. */
int ampe_surrogate_train(ampe_surrogate* s, long long n, const double* x,
                         const double* t, const ampe_train_params* tp,
                         float* loss_hist, ampe_stream_t stream);

/* Binary CALPHAD KKS over n independent cells with a per-cell temperature
 * (T device double[n]), the per-cell loop of
 * CALPHADequilibriumPhaseConcentrationsStrategy::
 * computePhaseConcentrationsOnPatch (source/
 * CALPHADequilibriumPhaseConcentrationsStrategy.cc:343-402) with the end
 * members evaluated from tab at each cell's T.  Without a surrogate the
 * Newton starts from (0.5, 0.5) (tests/testCALPHADbinaryKKS.cc:123-127).
 * With one, it starts from the surrogate's (cL, cS) for inputs (c, phi, T);
 * if after probe_its updates the residual exceeds probe_tol it restarts from
 * (0.5, 0.5).  Convergence: |residual| < p->newton_tol, at most
 * p->newton_max_its updates per start.  iters (device int[n], may be NULL):
 * updates per cell (both starts), -1 if not converged. */
typedef struct ampe_kks_batch_stats {
   long long iters;     /* Newton updates over all cells */
   long long max_iters; /* largest per-cell count */
   long long fallbacks; /* cells restarted from the default guess */
   long long failures;  /* cells not converged */
} ampe_kks_batch_stats;
int ampe_kks_batch(const ampe_model_params* p, const ampe_calphad_tdep* tab,
                   long long n, const double* c, const double* phi,
                   const double* T, const ampe_surrogate* guess, int probe_its,
                   double probe_tol, double* cl, double* cs, int* iters,
                   ampe_kks_batch_stats* stats, ampe_stream_t stream);

/* ampe_kks_batch without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as ampe_kks_batch, except that the statistics go to DEVICE memory
 * (stats_dev: one ampe_kks_batch_stats, may be NULL) and nothing is copied to
 * the host or synchronised.  The workgroup partials live in a file-scope
 * device array of 327,712 bytes per GPU (5 doubles for each of at most 8192
 * workgroups, plus 4 counters), held until the process ends; the batch grid
 * walks any n >= 0 with a grid stride, so no n exceeds it.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (different streams or
 * threads) race, so serialise them. */
int ampe_kks_batch_noalloc(const ampe_model_params* p, const ampe_calphad_tdep* tab,
                           long long n, const double* c, const double* phi,
                           const double* T, const ampe_surrogate* guess, int probe_its,
                           double probe_tol, double* cl, double* cs, int* iters,
                           ampe_kks_batch_stats* stats_dev, ampe_stream_t stream);

/* ------------------------------------------------------------- integrator */
/* BDF (order <= 2; imposed step, or CVODE's adaptive step and order with
 * solver.adaptive) Newton-Krylov integrator with SPGMR (modified or classical
 * Gram-Schmidt, left block preconditioner, difference-quotient Jv) driving
 * the kernels above -- the QuatIntegrator/CVODESolver callback loop
 * (source/QuatIntegrator.cc:2096-2233, 3134-3771;
 * source/samrai/CVODESolver.cc:150-247). */
typedef struct ampe_sim ampe_sim;
int ampe_sim_create(const ampe_model_params* p, const ampe_solver_params* s,
                    const ampe_box* boxes, int npatch,
                    const ampe_box* periodic_domain, ampe_sim** out);
int ampe_sim_destroy(ampe_sim* sim);
/* Host arrays of interior cells, patch after patch, i fastest; q holds qlen
 * component blocks.  c / q may be NULL when the model has no such field.
 * With norderp > 1 phi holds norderp component blocks (the phase RHS is then
 * MultiOrderRHSStrategy's, with one phase preconditioner block per order);
 * with the three-phase ternary model c holds two (and get_state's cl, cs the
 * liquid's and solid A's two species). */
int ampe_sim_set_state(ampe_sim* sim, const double* phi, const double* c,
                       const double* q);
int ampe_sim_get_state(ampe_sim* sim, double* phi, double* c, double* q,
                       double* cl, double* cs);
/* Advance nsteps accepted BDF steps; >0 if a Newton solve failed. */
int ampe_sim_advance(ampe_sim* sim, int nsteps, ampe_stats* stats);
/* Scalar temperature T(t) = T0 + dTdt t, held at T_target once reached
 * (ScalarTemperatureStrategy::getCurrentTemperature,
 * source/ScalarTemperatureStrategy.cc:56-72; Temperature { temperature,
 * dtemperaturedt, target_temperature } of the input deck), set at the time of
 * every step attempt as evaluateRHSFunction sets the temperature field at the
 * evaluation time (source/QuatIntegrator.cc:3191).  With tab (not NULL) the
 * CALPHAD phases are re-evaluated at that temperature
 * (ampe_calphad_at).  The model's own temperature applies before the first
 * step. */
int ampe_sim_set_temperature_ramp(ampe_sim* sim, double T0, double dTdt, double T_target,
                                  const ampe_calphad_tdep* tab);
/* Device time (ms) of the last ampe_sim_advance, from HIP events. */
int ampe_sim_last_time_ms(const ampe_sim* sim, double* ms);
/* Surrogate used when the solver's kks_guess is AMPE_GUESS_SURROGATE (the
 * integrator keeps the pointer; the object must outlive the sim).  A cell
 * whose previous-step guess has a KKS residual above switch_tol starts from
 * the surrogate's (cL, cS) for inputs (c, phi, T) instead; after probe_its
 * Newton updates a residual above probe_tol restarts it from the
 * previous-step guess.  switch_tol <= 0 starts every cell from the surrogate. */
int ampe_sim_set_surrogate(ampe_sim* sim, const ampe_surrogate* s,
                           int probe_its, double probe_tol, double switch_tol);

/* ------------------------------------------------- integrator callbacks */
/* The callbacks of CVODEAbstractFunctions (source/samrai/
 * CVODEAbstractFunctions.h:61-88) as QuatIntegrator implements them, for a
 * host that keeps its own CVODE (or other BDF / Newton-Krylov loop) and hands
 * the library caller-owned device state vectors.  A state vector holds the
 * sim's nc components (phi, c, q0..q3, as the model has them), each a ghosted
 * cell field of the sim's level, ampe_sim_vector_size() doubles in all; ghost
 * entries are scratch, filled by the callee.  Dots and WRMS norms over such
 * vectors: ampe_vec_dot / ampe_vec_wrms on ampe_sim_level() with ncomp = nc
 * (interior cells only; multiply dots by the cell volume for SAMRAI's
 * control-volume weighting).  The sim keeps the lagged face data, the phase
 * concentrations and the preconditioner operators between calls, as
 * QuatIntegrator does.  Work is enqueued on ampe_sim_stream(); run the
 * caller's own vector operations on that stream. */
int ampe_sim_vector_size(const ampe_sim* sim, long long* n, int* ncomp);
/* Borrowed handle of the sim's level (valid while the sim lives; not to be
 * passed to ampe_level_destroy). */
int ampe_sim_level(ampe_sim* sim, ampe_level** level);
int ampe_sim_stream(const ampe_sim* sim, ampe_stream_t* stream);
/* evaluateRHSFunction(t, y, ydot, fd_flag) (source/QuatIntegrator.cc
 * :3134-3298): ydot = f(t, y); fd_flag = 1 for the difference-quotient Jv
 * evaluations, which reuse the orientation face gradients and composition
 * face coefficients of the last fd_flag = 0 call (lag_quat_sidegrad).
 * Temperature at t when a ramp is set.  The per-cell KKS solve runs inside;
 * ampe_sim_kks_status() reports a failed cell. */
int ampe_rhs_eval(ampe_sim* sim, double t, double* y, double* ydot, int fd_flag);
/* CVSpgmrPrecondSet(t, y, fy, jok, jcur, gamma) (source/QuatIntegrator.cc
 * :3300-3376): coefficients of the phase, composition and orientation blocks
 * of I - gamma J at y (and the multigrid hierarchy when precond = MG);
 * *jcur = 1 (may be NULL). */
int ampe_precond_setup(ampe_sim* sim, double t, double* y, double gamma, int* jcur);
/* CVSpgmrPrecondSolve(t, y, fy, r, z, gamma, delta, lr) (source/
 * QuatIntegrator.cc:3666-3771, left preconditioning): z ~ P^-1 r with the
 * blocks of the last setup (its gamma; the gamma argument is accepted for
 * the CVODE signature).  0: every block's FAC loop reached the RMS tolerance
 * delta; 1: not (recoverable, as in the reference); < 0: error. */
int ampe_precond_solve(ampe_sim* sim, double t, const double* r, double* z, double gamma,
                       double delta);
/* applyProjection (source/QuatIntegrator.cc:3911-3965): corr = q/|q| - q on
 * the orientation components (0 elsewhere); err loses its component along
 * q/|q| there. */
int ampe_apply_projection(ampe_sim* sim, const double* y, double* corr, double* err);
/* resetRefPhaseConcentrations (source/QuatModel.cc:5218-5235): the last
 * phase concentrations become the KKS Newton guesses; call once per step. */
int ampe_sim_reset_ref_concentrations(ampe_sim* sim);
/* Synchronises the sim's stream; stats (may be NULL) receives the counters
 * of the calls so far; -1 if a KKS cell failed to converge. */
int ampe_sim_kks_status(ampe_sim* sim, ampe_stats* stats);

/* z = a x + b y over n doubles (N_VLinearSum; SAMRAIVectorReal::linearSum
 * as reached through Sundials_SAMRAIVector, source/samrai/
 * Sundials_SAMRAIVector.cc:106-203). */
int ampe_vec_axpby(long long n, double a, const double* x, double b, const double* y,
                   double* z, ampe_stream_t stream);


/* ------------------------------------------------------------ hierarchy */
/* A patch hierarchy of nlev levels, refinement ratio 2 (SAMRAI
 * PatchHierarchy as QuatModel builds it): level 0 covers the domain dom0
 * (cells), level l+1's boxes (in its own index space, aligned with level-l
 * cells) lie properly nested in level l; boxes[] lists npatch[0] boxes of
 * level 0, then level 1's, ...; dx0: level-0 spacing, equal in every
 * direction.  Fields on a hierarchy are composite arrays: per component,
 * every patch's ghosted block (ghost width 1), level by level; side fields
 * per direction, every patch's faces.  Each level is an ampe_level placed in
 * those arrays (ampe_hier_level), so every per-level entry point above runs
 * on it; ampe_hier_composite covers all patches (vector operations, KKS). */
typedef struct ampe_hier ampe_hier;
#define AMPE_MAX_LEVELS 4
int ampe_hier_create(int ndim, int nlev, const int* npatch, const ampe_box* boxes,
                     const int dom0[3], const double dx0[3], int periodic, ampe_hier** out);
int ampe_hier_destroy(ampe_hier* h);
int ampe_hier_sizes(const ampe_hier* h, long long* ncell_ghost, long long* ncell,
                    long long nside[3]);
int ampe_hier_level(ampe_hier* h, int l, ampe_level** level);
/* The hierarchy's levels: *nlev, npatch[l] (0 above the finest level) and
 * the boxes in ampe_hier_create's order (boxes may be NULL to query). */
int ampe_hier_boxes(const ampe_hier* h, int* nlev, int npatch[AMPE_MAX_LEVELS], ampe_box* boxes,
                    int max_boxes);
int ampe_hier_composite(ampe_hier* h, ampe_level** level);
/* Composite weights over ghosted cells: a cell's volume relative to a
 * level-0 cell, zero under a finer level and on ghosts (the control volumes
 * of QuatModel::computeVectorWeights, source/QuatModel.cc:4710-4750). */
int ampe_hier_weights(const ampe_hier* h, double* w, ampe_stream_t stream);
/* Ghost cells of every level, coarsest first: from the level (periodic
 * images; zero-slope images across the physical boundary), and on the
 * coarse-fine boundary from the next coarser level: cf = 0 leaves them,
 * 1 CONSTANT_REFINE (the coarse parent), 2 LINEAR_REFINE (quatlinrefine's
 * interpolation without normalisation) -- the refine schedules of
 * QuatIntegrator::fillScratch (source/QuatIntegrator.cc:2873-2957). */
int ampe_hier_ghost_fill(const ampe_hier* h, double* f, int ncomp, int cf,
                         ampe_stream_t stream);
/* Regrid (PFModel::Regrid, source/PFModel.cc:565-590, the refine schedule
 * of QuatModel::Regrid and initializeLevelData, source/QuatModel.cc:
 * 2733-2790, 2997-3004).
 * ampe_hier_regrid: a new hierarchy (the caller destroys it) with the old
 * level 0 and finer levels from the gradient detector, coarsest first: for
 * k = 0 .. max_levels-2 the tag field phi (one composite component on the
 * old hierarchy) is carried onto the new levels 0..k as ampe_hier_transfer
 * does, its ghosts filled by linear refinement, level k tagged as
 * ampe_tag_cells does (tol, the level's own spacing), and the tags
 * clustered as ampe_cluster_mask does (cp): over the level-0 domain for
 * k = 0, above per patch of level k inside the patch shrunk by one cell on
 * every side that is not a non-periodic physical boundary (proper nesting);
 * the clustered boxes refined by 2 form level k+1, and a level without tags
 * ends the hierarchy.  Synchronises the stream.
 * ampe_hier_transfer: interior cells of new-hierarchy field f_new (ncomp
 * components) from f_old on old_h: copied where a patch of the old
 * hierarchy's same level holds the cell, else LINEAR_REFINE of the new
 * level l-1 (quatlinrefine's interpolation, as the coarse-fine ghost fill),
 * coarsest level first; level 0 must be covered by the old level 0.
 * Ghosts are left as they are.  Synchronises the stream. */
int ampe_hier_regrid(const ampe_hier* old_h, const double* phi, const ampe_model_params* p0,
                     double tol, const ampe_cluster_params* cp, int max_levels, ampe_hier** out,
                     ampe_stream_t stream);
int ampe_hier_transfer(const ampe_hier* old_h, const ampe_hier* new_h, const double* f_old,
                       double* f_new, int ncomp, ampe_stream_t stream);
/* CONSERVATIVE_COARSEN of ncomp components from every finer level onto the
 * cells it covers, finest first (QuatIntegrator::coarsenData, source/
 * QuatIntegrator.cc:2289-2331); quatnorm: then q/|q| (QuatWeightedAverage,
 * at most four components). */
int ampe_hier_coarsen(const ampe_hier* h, double* f, int ncomp, int quatnorm,
                      ampe_stream_t stream);
/* Composite right-hand sides, ghosts of the inputs filled on every level:
 * each level's fused kernel, then the finer level's fluxes coarsened onto
 * the coarse-fine faces (PhaseRHSStrategyWithQ::evaluateRHS, source/
 * PhaseRHSStrategyWithQ.cc:90-149; QuatIntegrator::evaluateConcentrationRHS,
 * :2578-2690; QuatFACOps::evaluateRHS, source/QuatFACOps.cc:1861-1890).
 * p0 carries the level-0 spacing.  Only cells not under a finer level are
 * defined (the others are overwritten by coarsening in the reference). */
int ampe_hier_phase_rhs(const ampe_hier* h, const ampe_model_params* p0, const double* phi,
                        const double* q, const double* cl, const double* cs, double* rhs,
                        ampe_stream_t stream);
int ampe_hier_quat_face_norm2(const ampe_hier* h, const ampe_model_params* p0, const double* q,
                              double* n2, ampe_stream_t stream);
/* fc: side-field workspace (face coefficients) */
int ampe_hier_quat_rhs(const ampe_hier* h, const ampe_model_params* p0, const double* phi,
                       const double* q, const double* n2, double* fc, double* rhs,
                       ampe_stream_t stream);
/* KKS-form D_c, D_phi on every level's faces, then coarsened onto the
 * coarser faces (source/QuatIntegrator.cc:2395-2428). */
int ampe_hier_conc_coeffs(const ampe_hier* h, const ampe_model_params* p0, const double* phi,
                          const double* cl, const double* cs, double* Dc, double* Dphi,
                          ampe_stream_t stream);
int ampe_hier_conc_rhs(const ampe_hier* h, const ampe_model_params* p0, const double* c,
                       const double* phi, const double* Dc, const double* Dphi, double* rhs,
                       ampe_stream_t stream);

/* One block of the FAC preconditioner on the hierarchy.  kind 0: elliptic
 * A u = M div(D grad u) + C u (EllipticFACOps; Ewing's coarse-fine flux fix
 * when ewing != 0, the reference's default cf_discretization); kind 1: the
 * orientation block u + gamma S div(D grad(S u)) with S = sqrt(M_q) (cells,
 * ghosts filled) and D the face coefficients, constant-refined coarse-fine
 * ghosts (QuatFACOps stops on "Ewing" above level 0, source/QuatFACOps.cc:
 * 1093-1105, so the choice that runs is taken), red-black level smoothing
 * in the role of its hypre level solves. */
typedef struct ampe_hier_block {
   int kind;
   const double* D;
   double D_const;
   const double* C;
   double C_const;
   const double* M;
   double M_const;
   const double* S;
   double gamma;
   int ewing;
} ampe_hier_block;
typedef struct ampe_fac_params {
   int max_cycles;   /* FACPreconditioner max_cycles */
   int pre, post;    /* pre- and post-smoothing sweeps */
   int coarse_its;   /* coarse_solver_max_iterations ("redblack") */
   double coarse_tol; /* coarse_solver_tolerance (< 0: none) */
} ampe_fac_params;
/* FACPreconditioner::solveSystem / facCycle_Recursive (source/samrai/
 * FACPreconditioner.cc:267-474, 661-680) with EllipticFACOps' smoothing,
 * composite residual, restriction and CONSTANT_REFINE prolongation (source/
 * EllipticFACOps.cc:948-1241, 1480-1790): cycles from u (ghosted, ncomp
 * components: 1 for kind 0) until the weighted RMS composite residual is
 * below tol.  0: converged, 1: not (recoverable), < 0: error; *norm and
 * *cycles (host, may be NULL) report the last norm and the cycles.
 * Synchronises the stream once per cycle (the norm). */
int ampe_hier_fac_solve(ampe_hier* h, const ampe_hier_block* b, const ampe_fac_params* fp,
                        const double* f, double* u, int ncomp, double tol, double* norm,
                        int* cycles, ampe_stream_t stream);

/* The integrator on a hierarchy (QuatIntegrator over all levels of the
 * PatchHierarchy): state vectors are composite fields of nc components
 * (phase, composition, orientation) whose control volumes are the
 * hierarchy's weights; every right-hand-side evaluation coarsens the iterate
 * and fills the ghosts by linear refinement (setCoefficients / fillScratch,
 * source/QuatIntegrator.cc:2873-3083); the preconditioner blocks are FAC
 * cycles over the levels (ampe_hier_fac_solve; fac_max_cycles, fac_sweeps as
 * the coarse red-black iterations, fac_tol, one pre- and post-smoothing
 * sweep; quat_max_cycles, quat_sweeps for the orientation block).  Requires
 * precond redblack, precision fp64, the KKS composition form and the
 * previous-step KKS guess.  The hierarchy must outlive the sim. */
int ampe_sim_create_hier(const ampe_model_params* p, const ampe_solver_params* s,
                         ampe_hier* hier, ampe_sim** out);
/* host composite state (nc x the hierarchy's ghosted cells; ghost entries
 * ignored): coarsened, ghosts filled, initial phase concentrations */
int ampe_sim_set_state_composite(ampe_sim* sim, const double* y);
/* After a set_state, restart the integrator at time t (a regrid: the
 * reference resets CVODE at the current time with the previous step as its
 * first, QuatIntegrator::resetIntegrator / setSundialsOptions, source/
 * QuatIntegrator.cc:1571-1580, 1676-1730); h0 > 0 is the first step, else
 * the solver's own.  Works for single-level sims too. */
int ampe_sim_set_time(ampe_sim* sim, double t, double h0);
/* host composite state with the finer levels coarsened onto the cells they
 * cover (the output the reference coarsens after each step) */
int ampe_sim_get_state_composite(ampe_sim* sim, double* y);

#ifdef __cplusplus
}
#endif

#endif
