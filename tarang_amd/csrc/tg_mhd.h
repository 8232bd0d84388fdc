// tg_mhd.h -- launchers of the MHD pointwise stage kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "tg_kernels.h"

namespace tg {

hipError_t launch_mhd_elsasser(const Grid &g, double *const *V, double *const *W, double *const *Zp,
                               double *const *Zm, hipStream_t st);
// the same on the kz = 0 and kz = Nz/2 planes only
hipError_t launch_mhd_elsasser_planes(const Grid &g, double *const *V, double *const *W, double *const *Zp,
                                      double *const *Zm, hipStream_t st);
// NU <- (NU + NW)/2, NW <- (NU - NW)/2 in place (MHD::Elsasser_to_UB_nlin)
hipError_t launch_mhd_combine(const Grid &g, double *const *NU, double *const *NW, hipStream_t st);
// Fused x pass of an MHD RK stage (1..4): P[3i+j] = FT[Z-_j Z+_i] (work
// layout; the eight trace-shifted products, P_xx and P_yy less P_zz, P[8]
// not read), state V (u), W (b) in the FP64 boundary layout, accumulators
// accV, accW FP64 in the padded work layout; every stage writes the x-inverse
// of the next stage's Z- into U[0..2] and Z+ into U[3..5] (stage 4: of the
// next step's first stage).
struct XmhdArgs {
    void *P[9];
    double *V[3], *W[3];
    double *accV[3], *accW[3];
    void *U[6];
    const void *tw;
    Grid g;
    int stage;
    double dt, fh_u, f1_u, fh_b, f1_b;  // fh = -nu dt / 2, f1 = -nu dt (u: nu, b: eta)
    double scale;                       // 1 / (Nx Ny Nz)
    int *nonfinite;                     // stage 4: set to 1 on a non-finite state value
};

template <typename T>
hipError_t launch_x_mhd(const XmhdArgs &a, hipStream_t st);

}  // namespace tg
