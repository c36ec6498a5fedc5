// Patch hierarchy interface (see amr.hip).
#pragma once

#include "kernels.h"

namespace ampe {

struct Hier;

// FAC work arrays of one element type, grown on demand (amr.hip)
template <typename T>
struct FacWork {
   T *e = nullptr, *r = nullptr, *tmp = nullptr, *oflux = nullptr, *gsave = nullptr;
   long long cap = 0;
   void release()
   {
      T* all[] = {e, r, tmp, oflux, gsave};
      for (T* v : all) (void)hipFree(v);
      e = r = tmp = oflux = gsave = nullptr;
      cap = 0;
   }
};

// One preconditioner block of the FAC solve over the levels in the element
// type T: kind 0 elliptic (D on faces or D_const, C, M), kind 1 the
// orientation block (face coefficients D, S = sqrt(M_q), gamma), as
// ampe_hier_block.
template <typename T>
struct HBlock {
   int kind;
   const T* D;
   T D_const;
   const T* C;
   T C_const;
   const T* M;
   T M_const;
   const T* S;
   T gamma;
   int ewing;
};
// the hierarchy of a C-ABI handle
Hier* hier_of(ampe_hier* h);

int hier_create(int ndim, int nlev, const int* npatch, const ampe_box* boxes, const int dom0[3],
                const double dx0[3], int periodic, Hier** out);
void hier_destroy(Hier* H);
Level* hier_composite(Hier* H);
int hier_weight_sum(Hier* H, double* sum);
int hier_nlev(const Hier* H);
// coarsenData of an integrator state (phase, composition: CONSERVATIVE_COARSEN;
// orientation: QuatWeightedAverage) and its ghosts by LINEAR_REFINE
int hier_sim_sync_fill(Hier* H, double* y, int nc, int nord, int ic, int ncs, int iq, int qlen,
                       hipStream_t s);
int hier_ghost_fill(const Hier& H, double* f, int ncomp, int cf, hipStream_t s);
int hier_coarsen(const Hier& H, double* f, int ncomp, int quatnorm, hipStream_t s);
int hier_phase_rhs(const Hier& H, const ampe_model_params& p0, const double* phi,
                   const double* q, const double* cl, const double* cs, double* rhs,
                   hipStream_t s);
int hier_quat_face_norm2(const Hier& H, const ampe_model_params& p0, const double* q,
                         double* n2, hipStream_t s);
int hier_quat_rhs(const Hier& H, const ampe_model_params& p0, const double* phi, const double* q,
                  const double* n2, double* fc, double* rhs, hipStream_t s);
int hier_conc_coeffs(const Hier& H, const ampe_model_params& p0, const double* phi,
                     const double* cl, const double* cs, double* Dc, double* Dphi, hipStream_t s);
int hier_conc_rhs(const Hier& H, const ampe_model_params& p0, const double* c, const double* phi,
                  const double* Dc, const double* Dphi, double* rhs, hipStream_t s);
// the three-phase multi-order model: composite phase RHS of every order,
// the face diffusivities of the three phases and both species' EBS RHS
int hier_phase_rhs_tern3(const Hier& H, const ampe_model_params& p0, int norder,
                         const double* phi, const double* ph6, double* rhs, hipStream_t s);
int hier_phase_rhs_multiorder(const Hier& H, const ampe_model_params& p0, int norder,
                              const double* phi, double* rhs, hipStream_t s);
int hier_ebs3_coeffs(const Hier& H, const ampe_model_params& p0, int norder, const double* phi,
                     double* DL, double* DA, double* DB, hipStream_t s);
int hier_conc_rhs_ebs3(const Hier& H, const ampe_model_params& p0, const double* ph6,
                       const double* DL, const double* DA, const double* DB, double* rhs,
                       hipStream_t s);
// regrid: data of a new hierarchy from an old one, and new level boxes from
// the gradient detector (amr.hip)
int hier_transfer(const Hier& Ho, const Hier& Hn, const double* fo, double* fn, int ncomp,
                  hipStream_t s);
int hier_regrid(const Hier& Ho, const double* phi_o, const ampe_model_params& p0, double tol,
                const ampe_cluster_params& cp, int max_levels, Hier** out, hipStream_t s);
int hier_fac_solve(Hier& H, const ampe_hier_block& B, const ampe_fac_params& P, const double* f,
                   double* u, int ncomp, double tol, double* norm, int* cycles, hipStream_t s);
// the same block solve in FP32 (f, u, the coefficients and the cycle's work
// arrays), for the FP32 preconditioner over a hierarchy
int hier_fac_solve_f32(Hier& H, const HBlock<float>& B, const ampe_fac_params& P, const float* f,
                       float* u, int ncomp, double tol, double* norm, int* cycles, hipStream_t s);
// Per-level coefficient passes of the preconditioner setup: phase C and
// mobility, orientation face coefficients and sqrt(M_q), each level with its
// own spacing.
int hier_precond_coeffs(const Hier& H, const ampe_model_params& p0, double gamma,
                        const double* phi, const double* cl, const double* cs, const double* n2,
                        double* Mphi, double* Cphi, double* fc, double* sqrtm, hipStream_t s);

}  // namespace ampe
