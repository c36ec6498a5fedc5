// Per-cell KKS phase-concentration solves: the integrator's level-wide call
// and the batch solve over independent cells with a per-cell temperature.
#pragma once

#include "kernels.h"
#include "mlp.h"

namespace ampe {

// Initial-guess policy of the integrator's KKS: surrogate == nullptr starts
// every cell from (clr, csr) (the previous step's concentrations); otherwise a
// cell whose residual at (clr, csr) exceeds switch_tol starts from the
// surrogate, restarting from (clr, csr) when the residual after probe_its
// updates exceeds probe_tol.
struct KksGuess {
   const Surrogate* surrogate = nullptr;
   int probe_its = 3;
   double probe_tol = 1.0e-8;
   double switch_tol = 1.0e-2;  // previous-guess residual above which the surrogate starts
   // device work list of the cells that start from the surrogate (ncg
   // entries) and its counter (zero between calls)
   int* list = nullptr;
   unsigned* count = nullptr;
};

// Workgroups of the surrogate pass (each stages the network once and walks
// the work list with a grid stride).
constexpr int kKksSurrGrid = 256;

// KKS over all ghosted cells of a level.  launch_kks accumulates into device
// counters dev_acc[6] = {total iterations, max iterations, failure flag,
// calls, fallbacks, surrogate starts} without synchronising; launch_kks_sync returns max
// iterations (or -1 on a non-converged cell) and the total, synchronising s.
constexpr int kKksAcc = 6;
int launch_kks(const Level& L, const ampe_model_params& p, const double* phi,
               const double* c, const double* clr, const double* csr, double* cl,
               double* cs, const KksGuess& guess, long long* dev_acc, hipStream_t s);
int launch_kks_sync(const Level& L, const ampe_model_params& p, const double* phi,
                    const double* c, const double* clr, const double* csr, double* cl,
                    double* cs, int* max_iters, long long* total_iters, hipStream_t s);

int kks_batch(const ampe_model_params& p, const ampe_calphad_tdep& tab, long long n,
              const double* c, const double* phi, const double* T, const Surrogate* guess,
              int probe_its, double probe_tol, double* cl, double* cs, int* iters,
              ampe_kks_batch_stats* stats, hipStream_t s);
// The same without device allocation or host synchronisation: workgroup
// partials in a file-scope device array; statistics to device memory.
int kks_batch_noalloc(const ampe_model_params& p, const ampe_calphad_tdep& tab, long long n,
                      const double* c, const double* phi, const double* T,
                      const Surrogate* guess, int probe_its, double probe_tol, double* cl,
                      double* cs, int* iters, ampe_kks_batch_stats* stats_dev, hipStream_t s);

}  // namespace ampe
