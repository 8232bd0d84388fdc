// tg_chff.h -- launchers of the Chebyshev (ChFF) line kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "tg_kernels.h"

namespace tg {

// Helmholtz solves of one array: R -> U, per (ky, kz) column, lambda =
// Kperp^2 + lam_sup, wall values vp (x = 1) and vm (x = -1) per part (real,
// imaginary).  E = 2 for complex arrays, 1 for real arrays; X, Y: scratch of
// (ny * nzc * 2 * E) * (nx / 2) doubles each.
struct HelmArgs {
    const double *R;
    double *U;
    double *X, *Y;
    Grid g;
    double lam_sup;
    double vp[2], vm[2];
    int E;
    int pressure_switch;  // Kperp = 0 with a pressure: the linear profile
};

// Compute_pressure_Ux_chebyshev: P, V1 complex; homogeneous solutions real
struct InflArgs {
    double *P, *V1;
    const double *pp, *pm, *vxp, *vxm, *M;
    int nx, ny, nzc;
    double kf;
};

hipError_t launch_cheb_scale(const Grid &g, const double *src, double *dst, double w, hipStream_t st);
hipError_t launch_cheb_deriv(const Grid &g, const double *A, double *B, bool complex_, hipStream_t st);
hipError_t launch_helmholtz(const HelmArgs &a, hipStream_t st);
hipError_t launch_influence(const Grid &g, const double *vxp, const double *vxm, double *M, hipStream_t st);
hipError_t launch_influence_correct(const InflArgs &a, hipStream_t st);
hipError_t launch_real_product(const double *a, const double *b, double *c, size_t n, hipStream_t st);
hipError_t launch_add_deriv(const Grid &g, double *dst, const double *P, const double *D, int dir, bool assign,
                            hipStream_t st);
hipError_t launch_bdf1_rhs(const Grid &g, double *S, const double *V, const double *P, const double *D, int comp,
                           double dt, double nu, hipStream_t st);
hipError_t launch_scale_real(double *a, size_t n, double divisor, hipStream_t st);
hipError_t launch_fill_double(double *a, size_t n, double v, hipStream_t st);

}  // namespace tg
