// extern "C" entry points of libampe_hip.so (declared in include/ampe_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/ampe_hip.h"
#include "kks.h"
#include "mg.h"
#include "sim.h"

namespace ampe {
const char* last_error();
}

using ampe::Level;

struct ampe_level {
   Level* L;
   bool owned = true;  // false: a view of a sim's level (ampe_sim_level)
};
struct ampe_sim {
   ampe::Sim* S;
   ampe_level view{nullptr, false};
};
struct ampe_surrogate {
   ampe::Surrogate* S;
};

#define AMPE_REQUIRE(cond, msg)        \
   do {                                \
      if (!(cond)) {                   \
         ampe::set_error(msg);         \
         return -1;                    \
      }                                \
   } while (0)

static hipStream_t hs(ampe_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

static int check_params(const ampe_model_params* p, const Level& L)
{
   AMPE_REQUIRE(p, "null model params");
   AMPE_REQUIRE(p->ndim == L.ndim, "model ndim does not match the level");
   AMPE_REQUIRE(!p->evolve_quat || p->qlen == 4, "qlen must be 4");
   for (int d = 0; d < p->ndim; d++) AMPE_REQUIRE(p->dx[d] > 0.0, "dx must be positive");
   return 0;
}

namespace ampe {
// A non-owning handle on a level held elsewhere (the levels of a hierarchy).
ampe_level* level_view(Level* L) { return new ampe_level{L, false}; }
void level_view_free(ampe_level* v) { delete v; }
}  // namespace ampe

extern "C" {

const char* ampe_last_error(void) { return ampe::last_error(); }
int ampe_version(void) { return 1; }

int ampe_level_create(int ndim, const ampe_box* boxes, int npatch, int ghost,
                      const ampe_box* periodic_domain, ampe_level** out)
{
   AMPE_REQUIRE(out, "null output handle");
   Level* L = nullptr;
   if (ampe::level_create(ndim, boxes, npatch, ghost, periodic_domain, &L)) return -1;
   *out = new ampe_level{L};
   return 0;
}

int ampe_level_destroy(ampe_level* level)
{
   if (!level) return 0;
   AMPE_REQUIRE(level->owned, "ampe_level_destroy: the level belongs to a sim");
   ampe::level_destroy(level->L);
   delete level;
   return 0;
}

int ampe_level_sizes(const ampe_level* level, long long* ncell_ghost, long long* ncell,
                     long long nside[3])
{
   AMPE_REQUIRE(level, "null level");
   if (ncell_ghost) *ncell_ghost = level->L->ncg;
   if (ncell) *ncell = level->L->ncell;
   if (nside)
      for (int d = 0; d < 3; d++) nside[d] = level->L->nside[d];
   return 0;
}

int ampe_level_patch_offsets(const ampe_level* level, long long* cell_off, long long* side_off)
{
   AMPE_REQUIRE(level, "null level");
   const Level& L = *level->L;
   for (int p = 0; p < L.npatch; p++) {
      if (cell_off) cell_off[p] = L.hp[p].cell_off;
      if (side_off)
         for (int d = 0; d < 3; d++)
            side_off[3 * p + d] = L.side_dir_off[d] + L.hp[p].side_off[d];
   }
   return 0;
}

int ampe_level_patch_strides(const ampe_level* level, long long* sy, long long* sz)
{
   AMPE_REQUIRE(level, "null level");
   const Level& L = *level->L;
   for (int p = 0; p < L.npatch; p++) {
      if (sy) sy[p] = L.hp[p].sy;
      if (sz) sz[p] = L.hp[p].sz;
   }
   return 0;
}

int ampe_ghost_fill(const ampe_level* level, double* field, int ncomp, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && field, "null argument");
   return ampe::ghost_fill(*level->L, field, ncomp, hs(stream));
}

int ampe_quat_grad(const ampe_level* level, const ampe_model_params* p, const double* q,
                   double* gmod, double* face_norm2, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && q, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_quat_grad(*level->L, *p, q, gmod, face_norm2, hs(stream));
}

int ampe_phase_rhs(const ampe_level* level, const ampe_model_params* p, const double* phi,
                   const double* q, const double* cl, const double* cs, double* rhs,
                   ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   AMPE_REQUIRE(!p->evolve_quat || q, "orientation field required");
   AMPE_REQUIRE(p->flux_type != AMPE_FLUX_ANISOTROPIC || q, "anisotropy needs q");
   AMPE_REQUIRE(!p->with_conc || (cl && cs), "phase concentrations required");
   return ampe::launch_phase_rhs(*level->L, *p, phi, q, cl, cs, rhs, hs(stream));
}

int ampe_quat_rhs(const ampe_level* level, const ampe_model_params* p, const double* phi,
                  const double* q, const double* face_norm2, double* rhs, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && q && face_norm2 && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_quat_rhs(*level->L, *p, phi, q, face_norm2, rhs, hs(stream));
}

int ampe_quat_project(const ampe_level* level, const ampe_model_params* p, const double* q,
                      double* corr, double* err, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && q && corr && err, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_quat_project(*level->L, *p, q, corr, err, hs(stream));
}

int ampe_conc_coeffs(const ampe_level* level, const ampe_model_params* p, const double* phi,
                     const double* cl, const double* cs, double* Dc, double* Dphi,
                     ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && cl && cs && Dc && Dphi, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_conc_coeffs(*level->L, *p, phi, cl, cs, Dc, Dphi, hs(stream));
}

int ampe_quat_face_coefs(const ampe_level* level, const ampe_model_params* p, const double* phi,
                         const double* n2, double* fc, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && n2 && fc, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_quat_face_coef(*level->L, *p, phi, n2, fc, hs(stream));
}

int ampe_quat_rhs_fc(const ampe_level* level, const ampe_model_params* p, const double* phi,
                     const double* q, const double* fc, double* rhs, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && q && fc && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_quat_rhs_fc(*level->L, *p, phi, q, fc, rhs, hs(stream));
}

int ampe_grains_label(const ampe_level* level, const double* phase, double threshold,
                      int* labels, int* ngrains, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phase && labels, "null argument");
   AMPE_REQUIRE(threshold > 0.0, "the grain threshold must be positive");
   return ampe::grains_label(*level->L, phase, threshold, labels, ngrains, nullptr, hs(stream));
}

int ampe_kernel_timer_enable(int kernel, int on)
{
   return ampe::KernelTimer::enable(kernel, on != 0);
}

int ampe_kernel_timer_read(int kernel, double* ms, long long* launches, double* bytes)
{
   return ampe::KernelTimer::read(kernel, ms, launches, bytes);
}

int ampe_conc_coeffs_ebs(const ampe_level* level, const ampe_model_params* p,
                         const double* phi, double* DL, double* DA, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && DL && DA, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_ebs_coeffs(*level->L, *p, phi, DL, DA, hs(stream));
}

int ampe_conc_coeffs_ebs_mobility(const ampe_level* level, const ampe_model_params* p,
                                  const double* phi, const double* cl, const double* cs,
                                  double* DL, double* DA, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && cl && cs && DL && DA, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_ebs_coeffs_mob(*level->L, *p, phi, cl, cs, DL, DA, hs(stream));
}

int ampe_conc_rhs_ebs(const ampe_level* level, const ampe_model_params* p, const double* cl,
                      const double* cs, const double* DL, const double* DA, double* rhs,
                      ampe_stream_t stream)
{
   AMPE_REQUIRE(level && cl && cs && DL && DA && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_conc_rhs_ebs(*level->L, *p, cl, cs, DL, DA, rhs, hs(stream));
}

int ampe_phase_rhs_multiorder(const ampe_level* level, const ampe_model_params* p, int norder,
                              const double* phi, const double* mobility, double* rhs,
                              ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_phase_rhs_multiorder(*level->L, *p, norder, phi, mobility, rhs,
                                            hs(stream));
}

int ampe_kks_ternary3(const ampe_level* level, const ampe_model_params* p, const double* phi,
                      const double* c, double* ph, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && c && ph, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_kks_tern3(*level->L, *p, p->norderp, phi, c, ph, hs(stream));
}

int ampe_phase_rhs_ternary3(const ampe_level* level, const ampe_model_params* p,
                            const double* phi, const double* ph, const double* mobility,
                            double* rhs, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && ph && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_phase_rhs_tern3(*level->L, *p, p->norderp, phi, ph, mobility, rhs,
                                       hs(stream));
}

int ampe_ebs3_coeffs(const ampe_level* level, const ampe_model_params* p, const double* phi,
                     double* DL, double* DA, double* DB, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && DL && DA && DB, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_ebs3_coeffs(*level->L, *p, p->norderp, phi, DL, DA, DB, hs(stream));
}

int ampe_conc_rhs_ebs3(const ampe_level* level, const ampe_model_params* p, const double* ph,
                       const double* DL, const double* DA, const double* DB, double* rhs,
                       ampe_stream_t stream)
{
   AMPE_REQUIRE(level && ph && DL && DA && DB && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_conc_rhs_ebs3(*level->L, *p, ph, DL, DA, DB, rhs, hs(stream));
}

int ampe_project_phases(const ampe_level* level, int norder, const double* phi, double* corr,
                        double* err, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && corr && err, "null argument");
   return ampe::launch_project_phases(*level->L, norder, phi, corr, err, hs(stream));
}

int ampe_conc_rhs_ebs_4th(const ampe_level* level, const ampe_model_params* p, const double* cl,
                          const double* cs, const double* DL, const double* DA, double* rhs,
                          ampe_stream_t stream)
{
   AMPE_REQUIRE(level && cl && cs && DL && DA && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_conc_rhs_ebs_4th(*level->L, *p, cl, cs, DL, DA, rhs, hs(stream));
}

int ampe_conc_rhs_ebs_iso(const ampe_level* level, const ampe_model_params* p,
                          const double* phi_d, const double* cl, const double* cs,
                          double* rhs, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi_d && cl && cs && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_conc_rhs_ebs_iso(*level->L, *p, phi_d, cl, cs, rhs, hs(stream));
}

int ampe_conc_rhs(const ampe_level* level, const ampe_model_params* p, const double* c,
                  const double* phi, const double* Dc, const double* Dphi, double* rhs,
                  ampe_stream_t stream)
{
   AMPE_REQUIRE(level && c && phi && Dc && Dphi && rhs, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_conc_rhs(*level->L, *p, c, phi, Dc, Dphi, rhs, hs(stream));
}

int ampe_kks_solve(const ampe_level* level, const ampe_model_params* p, const double* phi,
                   const double* c, const double* cl_ref, const double* cs_ref, double* cl,
                   double* cs, int* max_iters, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && c && cl_ref && cs_ref && cl && cs, "null argument");
   if (check_params(p, *level->L)) return -1;
   int it = 0;
   if (ampe::launch_kks_sync(*level->L, *p, phi, c, cl_ref, cs_ref, cl, cs, &it, nullptr,
                             hs(stream)))
      return -1;
   if (max_iters) *max_iters = it;
   return it < 0 ? 1 : 0;
}

int ampe_tag_cells(const ampe_level* level, const ampe_model_params* p, const double* phi,
                   double tol, int* tags, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && tags, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_tag(*level->L, *p, phi, tol, tags, hs(stream));
}

int ampe_cluster_mask(int ndim, const ampe_box* domain, const unsigned char* mask,
                      const ampe_cluster_params* cp, ampe_box* boxes, int max_boxes,
                      int* nboxes)
{
   AMPE_REQUIRE(domain && mask && cp && boxes && nboxes, "null argument");
   return ampe::cluster_mask(ndim, *domain, mask, *cp, boxes, max_boxes, nboxes);
}

int ampe_cluster_tags(const ampe_level* level, const int* tags, const ampe_cluster_params* cp,
                      ampe_box* boxes, int max_boxes, int* nboxes, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && tags && cp && boxes && nboxes, "null argument");
   return ampe::cluster_tags(*level->L, tags, *cp, boxes, max_boxes, nboxes, hs(stream));
}

int ampe_fac_phase_coeffs(const ampe_level* level, const ampe_model_params* p, double gamma,
                          const double* phi, const double* M, double* C,
                          ampe_stream_t stream)
{
   AMPE_REQUIRE(level && phi && C, "null argument");
   if (check_params(p, *level->L)) return -1;
   return ampe::launch_phase_C(*level->L, *p, gamma, phi, M, C, hs(stream));
}

int ampe_phase_mobility(const ampe_level* level, const ampe_model_params* p, const double* cl,
                        const double* cs, double* M, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && M, "null argument");
   if (check_params(p, *level->L)) return -1;
   AMPE_REQUIRE(p->phi_mobility_type != AMPE_PHI_MOB_KIM || (cl && cs),
                "Kim mobility needs the phase concentrations");
   return ampe::launch_phase_mobility(*level->L, *p, cl, cs, M, hs(stream));
}

static ampe::EllOp<double> to_op(const ampe_elliptic_op* op)
{
   ampe::EllOp<double> o;
   o.D = op->D;
   o.Dc = op->D_const;
   o.C = op->C;
   o.Cc = op->C_const;
   o.M = op->M;
   o.Mc = op->M_const;
   return o;
}

int ampe_fac_smooth(const ampe_level* level, const ampe_elliptic_op* op, const double* rhs,
                    double* u, int num_sweeps, double tol, double* maxres,
                    ampe_stream_t stream)
{
   AMPE_REQUIRE(level && op && rhs && u, "null argument");
   const Level& L = *level->L;
   hipStream_t s = hs(stream);
   ampe::EllOp<double> o = to_op(op);
   bool want = tol >= 0.0 || maxres;
   double m = 0.0;
   if (ampe::ghost_fill(L, u, 1, s)) return -1;
   for (int sw = 0; sw < num_sweeps; sw++) {
      double mr = 0.0, mb = 0.0;
      if (ampe::ghost_fill(L, u, 1, s)) return -1;
      if (ampe::launch_rbgs<double>(L, o, op->dx, rhs, u, 0, want, s)) return -1;
      if (want && ampe::reduce_partials(L, L.ncell_tiles, 1, 1, &mr, s)) return -1;
      if (ampe::ghost_fill(L, u, 1, s)) return -1;
      if (ampe::launch_rbgs<double>(L, o, op->dx, rhs, u, 1, want, s)) return -1;
      if (want && ampe::reduce_partials(L, L.ncell_tiles, 1, 1, &mb, s)) return -1;
      if (ampe::ghost_fill(L, u, 1, s)) return -1;
      m = mr > mb ? mr : mb;
      if (tol >= 0.0 && m <= tol) break;
   }
   if (maxres) *maxres = m;
   return 0;
}

int ampe_quat_smooth(const ampe_level* level, const double* fc, const double* sm, double gamma,
                     const double dx[3], int ncomp, const double* rhs, double* z,
                     int num_sweeps, int fused, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && fc && sm && rhs && z && dx, "null argument");
   AMPE_REQUIRE(ncomp >= 1 && ncomp <= 4, "ncomp must be in [1, 4]");
   const Level& L = *level->L;
   hipStream_t s = hs(stream);
   if (!fused) {
      for (int sw = 0; sw < num_sweeps; sw++) {
         if (ampe::ghost_fill(L, z, ncomp, s)) return -1;
         if (ampe::launch_quat_rbgs<double>(L, fc, sm, gamma, dx, ncomp, rhs, z, 0, s)) return -1;
         if (ampe::ghost_fill(L, z, ncomp, s)) return -1;
         if (ampe::launch_quat_rbgs<double>(L, fc, sm, gamma, dx, ncomp, rhs, z, 1, s)) return -1;
      }
      return ampe::ghost_fill(L, z, ncomp, s);
   }
   double* t = nullptr;
   size_t bytes = sizeof(double) * (size_t)ncomp * (size_t)L.ncg;
   AMPE_HIP_CHECK(hipMalloc((void**)&t, bytes));
   // the second buffer starts as a copy, so ghosts without a gather source
   // hold z's values in both
   int rc = hipMemcpyAsync(t, z, bytes, hipMemcpyDeviceToDevice, s) == hipSuccess ? 0 : -1;
   double *cur = z, *oth = t;
   for (int sw = 0; sw < num_sweeps && !rc; sw++) {
      rc = ampe::ghost_fill(L, cur, ncomp, s);
      if (!rc)
         rc = ampe::launch_quat_rbgs_fused<double>(L, fc, sm, gamma, dx, ncomp, rhs, cur, oth, s);
      std::swap(cur, oth);
   }
   if (!rc) rc = ampe::ghost_fill(L, cur, ncomp, s);
   if (!rc && cur != z && hipMemcpyAsync(z, cur, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
      rc = -1;
   if (hipStreamSynchronize(s) != hipSuccess) rc = -1;
   (void)hipFree(t);
   if (rc) ampe::set_error(std::string("ampe_quat_smooth: ") + ampe::last_error());
   return rc;
}

int ampe_fac_residual(const ampe_level* level, const ampe_elliptic_op* op, const double* rhs,
                      double* u, double* res, double* rms, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && op && rhs && u && res, "null argument");
   const Level& L = *level->L;
   hipStream_t s = hs(stream);
   ampe::EllOp<double> o = to_op(op);
   if (ampe::ghost_fill(L, u, 1, s)) return -1;
   if (ampe::launch_residual<double>(L, o, op->dx, rhs, u, res, rms ? L.partial : nullptr, s))
      return -1;
   if (rms) {
      double ss = 0.0;
      if (ampe::reduce_partials(L, L.ncell_tiles, 1, 0, &ss, s)) return -1;
      *rms = std::sqrt(ss / (double)L.ncell);
   }
   return 0;
}

int ampe_fac_residual_restrict(const ampe_level* coarse, const ampe_level* fine,
                               const ampe_elliptic_op* op, const double* rhs, double* u,
                               double* res, double* b, int red_only, ampe_stream_t stream)
{
   AMPE_REQUIRE(coarse && fine && op && rhs && u && b, "null argument");
   AMPE_REQUIRE(ampe::is_coarsening(*coarse->L, *fine->L),
                "coarse level is not the fine level halved");
   const Level& F = *fine->L;
   hipStream_t s = hs(stream);
   ampe::EllOp<double> o = to_op(op);
   if (ampe::ghost_fill(F, u, 1, s)) return -1;
   if (red_only) {
      AMPE_REQUIRE(ampe::red_residual_ok(F, 1),
                   "red_only needs a periodic fine level with even, non-negative patch corners");
      return ampe::launch_ell_resrestrict<double>(*coarse->L, F, o, op->dx, rhs, u, b, s);
   }
   AMPE_REQUIRE(res, "null residual buffer");
   if (ampe::launch_residual<double>(F, o, op->dx, rhs, u, res, nullptr, s)) return -1;
   return ampe::mg_restrict_cell<double>(*coarse->L, F, 1, res, b, s);
}

int ampe_vec_dot(const ampe_level* level, int ncomp, const double* x, const double* y,
                 double* out, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && x && y && out, "null argument");
   return ampe::vec_dot(*level->L, ncomp, x, y, out, hs(stream));
}

int ampe_vec_wrms(const ampe_level* level, int ncomp, const double* x, const double* w,
                  double* out, ampe_stream_t stream)
{
   AMPE_REQUIRE(level && x && w && out, "null argument");
   double ss = 0.0;
   if (ampe::vec_sumsq_w(*level->L, ncomp, x, w, &ss, hs(stream))) return -1;
   *out = std::sqrt(ss / ((double)ncomp * level->L->ncell));
   return 0;
}

int ampe_sim_create(const ampe_model_params* p, const ampe_solver_params* s,
                    const ampe_box* boxes, int npatch, const ampe_box* periodic_domain,
                    ampe_sim** out)
{
   AMPE_REQUIRE(p && s && boxes && out, "null argument");
   AMPE_REQUIRE(p->ndim == 2 || p->ndim == 3, "ndim must be 2 or 3");
   AMPE_REQUIRE(s->dt > 0.0 && s->max_newton > 0, "bad solver parameters");
   ampe::Sim* S = nullptr;
   if (ampe::sim_create(*p, *s, boxes, npatch, periodic_domain, &S)) return -1;
   *out = new ampe_sim{S};
   (*out)->view.L = S->L;
   return 0;
}

int ampe_sim_create_hier(const ampe_model_params* p, const ampe_solver_params* s,
                         ampe_hier* hier, ampe_sim** out)
{
   AMPE_REQUIRE(p && s && hier && out, "null argument");
   ampe::Sim* S = nullptr;
   if (ampe::sim_create_hier(*p, *s, ampe::hier_of(hier), &S)) return -1;
   *out = new ampe_sim{S};
   (*out)->view.L = S->L;
   return 0;
}

int ampe_sim_set_state_composite(ampe_sim* sim, const double* y)
{
   AMPE_REQUIRE(sim && y, "null argument");
   return ampe::sim_set_state_composite(sim->S, y);
}

int ampe_sim_set_time(ampe_sim* sim, double t, double h0)
{
   AMPE_REQUIRE(sim, "null argument");
   return ampe::sim_set_time(sim->S, t, h0);
}

int ampe_sim_get_state_composite(ampe_sim* sim, double* y)
{
   AMPE_REQUIRE(sim && y, "null argument");
   return ampe::sim_get_state_composite(sim->S, y);
}

int ampe_sim_vector_size(const ampe_sim* sim, long long* n, int* ncomp)
{
   AMPE_REQUIRE(sim, "null argument");
   if (n) *n = sim->S->nv;
   if (ncomp) *ncomp = sim->S->nc;
   return 0;
}

int ampe_sim_level(ampe_sim* sim, ampe_level** level)
{
   AMPE_REQUIRE(sim && level, "null argument");
   *level = &sim->view;
   return 0;
}

int ampe_sim_stream(const ampe_sim* sim, ampe_stream_t* stream)
{
   AMPE_REQUIRE(sim && stream, "null argument");
   *stream = reinterpret_cast<ampe_stream_t>(sim->S->stream);
   return 0;
}

int ampe_rhs_eval(ampe_sim* sim, double t, double* y, double* ydot, int fd_flag)
{
   AMPE_REQUIRE(sim && y && ydot, "null argument");
   return ampe::sim_cb_rhs(sim->S, t, y, ydot, fd_flag);
}

int ampe_precond_setup(ampe_sim* sim, double t, double* y, double gamma, int* jcur)
{
   AMPE_REQUIRE(sim && y, "null argument");
   AMPE_REQUIRE(gamma > 0.0, "gamma must be positive");
   if (ampe::sim_cb_psetup(sim->S, t, y, gamma)) return -1;
   if (jcur) *jcur = 1;
   return 0;
}

int ampe_precond_solve(ampe_sim* sim, double t, const double* r, double* z, double gamma,
                       double delta)
{
   // the block operators were built by the last setup with its gamma, which
   // the solve uses (CVODE reuses a setup while gamma stays within 30 %)
   (void)t;
   (void)gamma;
   AMPE_REQUIRE(sim && r && z, "null argument");
   return ampe::sim_cb_psolve(sim->S, r, z, delta);
}

int ampe_apply_projection(ampe_sim* sim, const double* y, double* corr, double* err)
{
   AMPE_REQUIRE(sim && y && corr && err, "null argument");
   return ampe::sim_cb_project(sim->S, y, corr, err);
}

int ampe_sim_reset_ref_concentrations(ampe_sim* sim)
{
   AMPE_REQUIRE(sim, "null argument");
   return ampe::sim_cb_reset_ref(sim->S);
}

int ampe_sim_kks_status(ampe_sim* sim, ampe_stats* stats)
{
   AMPE_REQUIRE(sim, "null argument");
   int rc = ampe::sim_cb_kks_check(sim->S);
   if (stats) *stats = sim->S->st;
   return rc;
}

int ampe_vec_axpby(long long n, double a, const double* x, double b, const double* y,
                   double* z, ampe_stream_t stream)
{
   AMPE_REQUIRE(n >= 0, "negative n");
   AMPE_REQUIRE(n == 0 || (x && y && z), "null argument");
   if (n == 0) return 0;
   return ampe::vec_axpby(n, a, x, b, y, z, hs(stream));
}

int ampe_sim_destroy(ampe_sim* sim)
{
   if (!sim) return 0;
   ampe::sim_destroy(sim->S);
   delete sim;
   return 0;
}

int ampe_sim_set_temperature_ramp(ampe_sim* sim, double T0, double dTdt, double T_target,
                                  const ampe_calphad_tdep* tab)
{
   AMPE_REQUIRE(sim, "null argument");
   return ampe::sim_set_temperature_ramp(sim->S, T0, dTdt, T_target, tab);
}

int ampe_sim_set_state(ampe_sim* sim, const double* phi, const double* c, const double* q)
{
   AMPE_REQUIRE(sim && phi, "null argument");
   return ampe::sim_set_state(sim->S, phi, c, q);
}

int ampe_sim_get_state(ampe_sim* sim, double* phi, double* c, double* q, double* cl,
                       double* cs)
{
   AMPE_REQUIRE(sim, "null argument");
   return ampe::sim_get_state(sim->S, phi, c, q, cl, cs);
}

int ampe_sim_advance(ampe_sim* sim, int nsteps, ampe_stats* stats)
{
   AMPE_REQUIRE(sim && nsteps >= 0, "bad argument");
   return ampe::sim_advance(sim->S, nsteps, stats);
}

int ampe_sim_last_time_ms(const ampe_sim* sim, double* ms)
{
   AMPE_REQUIRE(sim && ms, "null argument");
   *ms = sim->S->last_ms;
   return 0;
}

int ampe_coarsen_cell(const ampe_level* coarse, const ampe_level* fine, int ncomp,
                      const double* f, double* c, ampe_stream_t stream)
{
   AMPE_REQUIRE(coarse && fine && f && c && ncomp >= 1, "bad argument");
   AMPE_REQUIRE(ampe::is_coarsening(*coarse->L, *fine->L),
                "coarse level is not the fine level halved");
   return ampe::mg_restrict_cell<double>(*coarse->L, *fine->L, ncomp, f, c, hs(stream));
}

int ampe_quat_coarsen(const ampe_level* coarse, const ampe_level* fine, const double* dx_fine,
                      int qlen, const double* qf, double* qc, ampe_stream_t stream)
{
   AMPE_REQUIRE(coarse && fine && dx_fine && qf && qc, "bad argument");
   AMPE_REQUIRE(ampe::is_coarsening(*coarse->L, *fine->L),
                "coarse level is not the fine level halved");
   return ampe::quat_coarsen(*coarse->L, *fine->L, dx_fine, qlen, qf, qc, hs(stream));
}

int ampe_linear_refine(const ampe_level* fine, const ampe_level* coarse, const double* dx_fine,
                       int ncomp, const double* c, double* f, ampe_stream_t stream)
{
   AMPE_REQUIRE(coarse && fine && dx_fine && c && f, "bad argument");
   AMPE_REQUIRE(ampe::is_coarsening(*coarse->L, *fine->L),
                "coarse level is not the fine level halved");
   return ampe::linear_refine(*fine->L, *coarse->L, dx_fine, ncomp, c, f, hs(stream));
}

int ampe_quat_refine(const ampe_level* fine, const ampe_level* coarse, const double* dx_fine,
                     int qlen, const double* qc, double* qf, ampe_stream_t stream)
{
   AMPE_REQUIRE(coarse && fine && dx_fine && qf && qc, "bad argument");
   AMPE_REQUIRE(ampe::is_coarsening(*coarse->L, *fine->L),
                "coarse level is not the fine level halved");
   return ampe::quat_refine(*fine->L, *coarse->L, dx_fine, qlen, qc, qf, hs(stream));
}

int ampe_coarsen_side(const ampe_level* coarse, const ampe_level* fine, const double* f,
                      double* c, ampe_stream_t stream)
{
   AMPE_REQUIRE(coarse && fine && f && c, "bad argument");
   AMPE_REQUIRE(ampe::is_coarsening(*coarse->L, *fine->L),
                "coarse level is not the fine level halved");
   return ampe::mg_restrict_side<double>(*coarse->L, *fine->L, f, c, hs(stream));
}

int ampe_refine_add(const ampe_level* fine, const ampe_level* coarse, int ncomp, int order,
                    const double* c, double* f, ampe_stream_t stream)
{
   AMPE_REQUIRE(coarse && fine && f && c && ncomp >= 1, "bad argument");
   AMPE_REQUIRE(order == 0 || order == 1, "order must be 0 or 1");
   AMPE_REQUIRE(ampe::is_coarsening(*coarse->L, *fine->L),
                "coarse level is not the fine level halved");
   return ampe::mg_prolong_add<double>(*fine->L, *coarse->L, ncomp, order, c, f, hs(stream));
}

int ampe_quat_residual_restrict(const ampe_level* coarse, const ampe_level* fine,
                                const double* fc, const double* sm, double gamma,
                                const double dx[3], int ncomp, const double* rhs, double* z,
                                double* res, double* b, int red_only, ampe_stream_t stream)
{
   AMPE_REQUIRE(coarse && fine && fc && sm && rhs && z && b && dx, "null argument");
   AMPE_REQUIRE(ncomp >= 1 && ncomp <= 4, "ncomp must be in [1, 4]");
   AMPE_REQUIRE(ampe::is_coarsening(*coarse->L, *fine->L),
                "coarse level is not the fine level halved");
   const Level& F = *fine->L;
   hipStream_t s = hs(stream);
   if (ampe::ghost_fill(F, z, ncomp, s)) return -1;
   if (red_only) {
      AMPE_REQUIRE(ampe::red_residual_ok(F, ncomp),
                   "red_only needs a periodic fine level with even, non-negative patch "
                   "corners and one, two or four components");
      return ampe::launch_quat_resrestrict<double>(*coarse->L, F, fc, sm, gamma, dx, ncomp, rhs,
                                                   z, b, s);
   }
   AMPE_REQUIRE(res, "null residual buffer");
   if (ampe::launch_quat_residual<double>(F, fc, sm, gamma, dx, ncomp, rhs, z, res, nullptr, s))
      return -1;
   return ampe::mg_restrict_cell<double>(*coarse->L, F, ncomp, res, b, s);
}

int ampe_sim_set_surrogate(ampe_sim* sim, const ampe_surrogate* s, int probe_its,
                           double probe_tol, double switch_tol)
{
   AMPE_REQUIRE(sim, "null argument");
   AMPE_REQUIRE(probe_its >= 0, "probe_its must be >= 0");
   if (s) {
      const ampe::Surrogate& m = *s->S;
      AMPE_REQUIRE(m.nin == 3 && m.nout == 2,
                   "the KKS surrogate maps (c, phi, T) -> (cL, cS): n_in 3, n_out 2");
   }
   if (s && !sim->S->guess.list) {
      ampe::KksGuess& g = sim->S->guess;
      hipError_t e = hipMalloc(&g.list, sizeof(int) * (size_t)sim->S->L->ncg);
      if (e == hipSuccess) e = hipMalloc(&g.count, sizeof(unsigned));
      if (e == hipSuccess) e = hipMemset(g.count, 0, sizeof(unsigned));
      if (e != hipSuccess) {
         (void)hipFree(g.list);
         (void)hipFree(g.count);
         g.list = nullptr;
         g.count = nullptr;
         ampe::set_error(std::string("ampe_sim_set_surrogate: ") + hipGetErrorString(e));
         return -1;
      }
   }
   sim->S->guess.surrogate = s ? s->S : nullptr;
   sim->S->guess.probe_its = probe_its;
   sim->S->guess.probe_tol = probe_tol;
   sim->S->guess.switch_tol = switch_tol;
   return 0;
}

int ampe_calphad_at(const ampe_calphad_tdep* tab, double T, ampe_calphad_phase out[2])
{
   AMPE_REQUIRE(tab && out, "null argument");
   AMPE_REQUIRE(T > 0.0, "temperature must be positive");
   for (int ph = 0; ph < 2; ph++)
      for (int sp = 0; sp < 2; sp++)
         AMPE_REQUIRE(tab->g[ph][sp].nrange >= 1 && tab->g[ph][sp].nrange <= 4,
                      "gfun nrange must be in [1, 4]");
   ampe::calphad_at(*tab, T, out);
   return 0;
}

int ampe_surrogate_create(int n_in, int n_hidden, int n_out, unsigned long long seed,
                          ampe_surrogate** out)
{
   AMPE_REQUIRE(out, "null output handle");
   ampe::Surrogate* S = nullptr;
   if (ampe::surrogate_create(n_in, n_hidden, n_out, seed, &S)) return -1;
   *out = new ampe_surrogate{S};
   return 0;
}

int ampe_surrogate_destroy(ampe_surrogate* s)
{
   if (!s) return 0;
   ampe::surrogate_destroy(s->S);
   delete s;
   return 0;
}

int ampe_surrogate_nparams(const ampe_surrogate* s, long long* n)
{
   AMPE_REQUIRE(s && n, "null argument");
   *n = s->S->nparam;
   return 0;
}

int ampe_surrogate_set_params(ampe_surrogate* s, const float* params)
{
   AMPE_REQUIRE(s && params, "null argument");
   return ampe::surrogate_set_params(s->S, params);
}

int ampe_surrogate_get_params(const ampe_surrogate* s, float* params)
{
   AMPE_REQUIRE(s && params, "null argument");
   return ampe::surrogate_get_params(s->S, params);
}

int ampe_surrogate_set_scaling(ampe_surrogate* s, const double* in_shift, const double* in_scale,
                               const double* out_shift, const double* out_scale)
{
   AMPE_REQUIRE(s && in_shift && in_scale && out_shift && out_scale, "null argument");
   ampe::Surrogate& m = *s->S;
   for (int i = 0; i < m.nin; i++) {
      AMPE_REQUIRE(std::isfinite(in_scale[i]) && in_scale[i] != 0.0, "bad input scale");
      m.in_shift[i] = (float)in_shift[i];
      m.in_scale[i] = (float)in_scale[i];
   }
   for (int i = 0; i < m.nout; i++) {
      AMPE_REQUIRE(std::isfinite(out_scale[i]) && out_scale[i] != 0.0, "bad output scale");
      m.out_shift[i] = out_shift[i];
      m.out_scale[i] = out_scale[i];
   }
   return 0;
}

int ampe_surrogate_forward(const ampe_surrogate* s, long long n, const float* x, float* y,
                           ampe_stream_t stream)
{
   AMPE_REQUIRE(s && (n == 0 || (x && y)), "null argument");
   AMPE_REQUIRE(n >= 0, "negative n");
   return ampe::surrogate_forward(s->S, n, x, y, hs(stream));
}

int ampe_surrogate_loss_grad(ampe_surrogate* s, long long n, const float* x, const float* t,
                             float* grad, double* loss, ampe_stream_t stream)
{
   AMPE_REQUIRE(s && x && t && grad, "null argument");
   return ampe::surrogate_loss_grad(s->S, n, x, t, grad, loss, hs(stream));
}

int ampe_surrogate_train(ampe_surrogate* s, long long n, const double* x, const double* t,
                         const ampe_train_params* tp, float* loss_hist, ampe_stream_t stream)
{
   AMPE_REQUIRE(s && x && t && tp, "null argument");
   AMPE_REQUIRE(tp->lr0 > 0.0 && tp->lr1 > 0.0, "learning rates must be positive");
   AMPE_REQUIRE(tp->beta1 >= 0.0 && tp->beta1 < 1.0 && tp->beta2 >= 0.0 && tp->beta2 < 1.0,
                "Adam betas must be in [0, 1)");
   return ampe::surrogate_train(s->S, n, x, t, *tp, loss_hist, hs(stream));
}

int ampe_kks_batch(const ampe_model_params* p, const ampe_calphad_tdep* tab, long long n,
                   const double* c, const double* phi, const double* T,
                   const ampe_surrogate* guess, int probe_its, double probe_tol, double* cl,
                   double* cs, int* iters, ampe_kks_batch_stats* stats, ampe_stream_t stream)
{
   AMPE_REQUIRE(p && tab, "null argument");
   AMPE_REQUIRE(n >= 0, "negative n");
   AMPE_REQUIRE(n == 0 || (c && phi && T && cl && cs), "null argument");
   AMPE_REQUIRE(!guess || probe_its >= 0, "probe_its must be >= 0");
   return ampe::kks_batch(*p, *tab, n, c, phi, T, guess ? guess->S : nullptr, probe_its,
                          probe_tol, cl, cs, iters, stats, hs(stream));
}

int ampe_kks_batch_noalloc(const ampe_model_params* p, const ampe_calphad_tdep* tab,
                           long long n, const double* c, const double* phi, const double* T,
                           const ampe_surrogate* guess, int probe_its, double probe_tol,
                           double* cl, double* cs, int* iters, ampe_kks_batch_stats* stats_dev,
                           ampe_stream_t stream)
{
   AMPE_REQUIRE(p && tab, "null argument");
   AMPE_REQUIRE(n >= 0, "negative n");
   AMPE_REQUIRE(n == 0 || (c && phi && T && cl && cs), "null argument");
   AMPE_REQUIRE(!guess || probe_its >= 0, "probe_its must be >= 0");
   return ampe::kks_batch_noalloc(*p, *tab, n, c, phi, T, guess ? guess->S : nullptr,
                                  probe_its, probe_tol, cl, cs, iters, stats_dev, hs(stream));
}

}  // extern "C"
