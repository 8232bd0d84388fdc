// tg_sff.h -- SFF basis (sin/cos along x, Fourier along y, z) kernels and the
// Rayleigh-Benard stage pipeline on it.
#pragma once
#include "tg_kernels.h"

namespace tg {

// Parity bits: bit c set = field c is a sine ('S') field along x, clear =
// cosine ('C').  The RBC state (u_x, u_y, u_z, theta) is S, C, C, S.
constexpr int kRbcStateParity = 0x9;

// x-inverse (sin/cos) of FP64 spectral fields into T work fields.
//   flags: 1 = dealias on read, 4 = only the kz = 0 and kz = Nz/2 columns.
template <typename T>
hipError_t launch_xinv_sc(const Grid &g, double *const *V, int ncomp, int parity, void *const *U, int flags,
                          const void *twx, const void *twq, hipStream_t st);

// x-forward (sin/cos) of T work fields into FP64 spectral fields, scaled.
template <typename T>
hipError_t launch_xfwd_sc(const Grid &g, void *const *W, int ncomp, int parity, double *const *out, double scale,
                          const void *twx, const void *twq, hipStream_t st);

// z pass of the RBC nonlinear terms: 4 lines in (u_x, u_y, u_z, theta), 9
// products out: xx, yy, zz, xy, yz, zx, x theta, y theta, z theta.
template <typename T>
hipError_t launch_z_rbc(const Grid &g, void *const *U, void *const *P, const void *twm, const void *twn,
                        hipStream_t st, bool shift = false);

struct RbcParams {
    int stage;            // 1..4; 0 = nonlinear terms (or right-hand sides) only
    int write_rhs;        // stage 0: force, projection and signs of Compute_rhs
    double dt;
    double fh_u, f1_u;    // -nu dt/2, -nu dt
    double fh_t, f1_t;    // -kappa dt/2, -kappa dt
    double scale;         // 1/(2 Nx Ny Nz)
    int *nonfinite;       // stage 4: set to 1 if the new state holds a
                          // non-finite value (null: not checked)
    int trace_free;       // velocity products trace-shifted (xx - zz, yy - zz,
                          // zz not formed): an RK stage
};

// Fused x pass of an RBC stage: sin/cos x-forward of the 9 products, N and
// N_T, buoyancy, SFF projection, Lawson RK4 stage for u (nu) and theta
// (kappa), next stage input (dealiased) x-inverted into U, its undealiased
// u_x and theta into f_out (the force of the next stage).
//   V[0..3]: state u_x, u_y, u_z, theta (FP64 boundary layout)
//   acc[0..3]: accumulators (FP64 padded work layout)
//   f_in[0..1]: theta, u_x of this stage's input (boundary layout)
//   out[0..3]: stage 0 outputs N_x, N_y, N_z, N_T
template <typename T>
hipError_t launch_x_rbc(const Grid &g, const RbcParams &rp, void *const *P, double *const *V, double *const *acc,
                        void *const *U, const double *const *f_in, double *const *f_out, double *const *out,
                        const void *twx, const void *twq, hipStream_t st);

// Reality conditions on SFF fields (SFF_pencil_basic.cc:157-189): mode 1
// strong (A(lx,ky,0) = conj A(lx,-ky,0) for -Ny/2 < ky <= 0, kz = Nz/2 plane
// zeroed), mode 2 weak (kz = Nz/2 plane zeroed).
hipError_t launch_sff_reality(const Grid &g, double *const *V, int ncomp, int mode, hipStream_t st);

}  // namespace tg
