// Control-variate kernels on the uniform grid (models without jumps):
// 'diag' dim 1..4, 'indep' Double-GBM dim 1..4, Heston.
#include "cv_mlp.h"

namespace sdemc {

int launch_cv_uniform(const sdemc_job& sj, const CvLaunch& a) {
  KJob job;
  make_kjob(sj, job);
  const int64_t nc = (int64_t)((job.n_paths + WAVE - 1) / WAVE);
  const int d = sj.model.dim;
  switch (sj.model.noise) {
    case SDEMC_NOISE_HESTON:
      return d == 2 ? launch_cv_uniform_d<2, FAM_HESTON>(job, a, nc) : SDEMC_ERR_UNSUPPORTED;
    case SDEMC_NOISE_DIAG:
      switch (d) {
        case 1: return launch_cv_uniform_d<1, FAM_MERTON>(job, a, nc);
        case 2: return launch_cv_uniform_d<2, FAM_MERTON>(job, a, nc);
        case 3: return launch_cv_uniform_d<3, FAM_MERTON>(job, a, nc);
        case 4: return launch_cv_uniform_d<4, FAM_MERTON>(job, a, nc);
        default: return SDEMC_ERR_UNSUPPORTED;
      }
    case SDEMC_NOISE_INDEP_FULL:
      switch (d) {
        case 1: return launch_cv_uniform_d<1, FAM_DOUBLE>(job, a, nc);
        case 2: return launch_cv_uniform_d<2, FAM_DOUBLE>(job, a, nc);
        case 3: return launch_cv_uniform_d<3, FAM_DOUBLE>(job, a, nc);
        case 4: return launch_cv_uniform_d<4, FAM_DOUBLE>(job, a, nc);
        default: return SDEMC_ERR_UNSUPPORTED;
      }
    default: return SDEMC_ERR_UNSUPPORTED;
  }
}

}  // namespace sdemc
