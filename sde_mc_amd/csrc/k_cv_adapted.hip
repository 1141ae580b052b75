// Control-variate kernels on the jump-adapted grid: lognormal or Levy jumps
// on 'diag' or 'indep' (common small-jump Brownian motion) models, dim 1..4.
#include "cv_mlp.h"

namespace sdemc {

int launch_cv_adapted(const sdemc_job& sj, const CvLaunch& a) {
  KJob job;
  make_kjob(sj, job);
  const int64_t nc = (int64_t)((job.n_paths + WAVE - 1) / WAVE);
  const int d = sj.model.dim;
  switch (sj.model.noise) {
    case SDEMC_NOISE_DIAG:
      switch (d) {
        case 1: return launch_cv_adapted_d<1, FAM_MERTON>(job, a, nc);
        case 2: return launch_cv_adapted_d<2, FAM_MERTON>(job, a, nc);
        case 3: return launch_cv_adapted_d<3, FAM_MERTON>(job, a, nc);
        case 4: return launch_cv_adapted_d<4, FAM_MERTON>(job, a, nc);
        default: return SDEMC_ERR_UNSUPPORTED;
      }
    case SDEMC_NOISE_INDEP_COMMON:
      switch (d) {
        case 1: return launch_cv_adapted_d<1, FAM_LEVY>(job, a, nc);
        case 2: return launch_cv_adapted_d<2, FAM_LEVY>(job, a, nc);
        case 3: return launch_cv_adapted_d<3, FAM_LEVY>(job, a, nc);
        case 4: return launch_cv_adapted_d<4, FAM_LEVY>(job, a, nc);
        default: return SDEMC_ERR_UNSUPPORTED;
      }
    default: return SDEMC_ERR_UNSUPPORTED;
  }
}

}  // namespace sdemc
