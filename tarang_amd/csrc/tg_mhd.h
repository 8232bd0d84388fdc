// tg_mhd.h -- launchers of the MHD pointwise stage kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "tg_kernels.h"

namespace tg {

hipError_t launch_mhd_elsasser(const Grid &g, double *const *V, double *const *W, double *const *Zp,
                               double *const *Zm, hipStream_t st);
// NU <- (NU + NW)/2, NW <- (NU - NW)/2 in place (MHD::Elsasser_to_UB_nlin)
hipError_t launch_mhd_combine(const Grid &g, double *const *NU, double *const *NW, hipStream_t st);
hipError_t launch_mhd_rk(const Grid &g, int stage, double dt, double nu, double eta, double *const *V,
                         double *const *W, double *const *accV, double *const *accW, const double *const *Np,
                         const double *const *Nm, double *const *Zp, double *const *Zm, hipStream_t st,
                         int *nonfinite = nullptr);

}  // namespace tg
