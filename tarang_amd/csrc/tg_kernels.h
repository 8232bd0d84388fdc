// tg_kernels.h -- kernel argument blocks and host-side launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

// Grid description shared by every kernel (reference FFF basis conventions:
// kx = lx <= Nx/2 ? lx : lx - Nx, likewise ky; kz = lz; K_d = k_d*kfactor[d]).
struct Grid {
    int nx, ny, nz, nzc;  // nzc = nz/2 + 1 (kz stride of FP64 state fields)
    double kf[3];
    int dealias;          // 0 none, 1 reference corner box, 2 two-thirds rule
    int nzp;              // kz stride of work fields and the RK accumulator:
                          // nzc rounded up to a multiple of 128 bytes
    int basis;            // 0 FFF; 1 SFF (sin/cos along x: kx = lx, K_x = lx*kf[0])
};

// Forcing band slots: modes with |kx|,|ky| <= kb and 0 <= kz <= kb.
struct Band {
    int kind;             // 0 none, 1 constant energy supply alpha-forcing, 2 random band forcing
    int kb;               // slot box half-width
    double r_in, r_out, eps;
    // kind 2: per-mode supplies (e_k, h_k) of shells 0..ceil(r_out), random
    // interval and the seed of the phases
    const double *shell_eh;
    double tau;
    unsigned long long seed;
};

__host__ __device__ inline int band_nslots(int kb) { return (2 * kb + 1) * (2 * kb + 1) * (kb + 1); }

// Launch helpers (defined in tg_kernels.hip).  All return hipError_t.
// T is float or double (transform precision).
template <typename T>
hipError_t launch_colfft(const Grid &g, int axis, int dir, void *const *fields, int nfields,
                         const void *tw, hipStream_t st);

// x-inverse of an FP64 spectral field set into the T work layout.
//   flags: 1 = dealias on read, 2 = write band slots, 4 = reality patch
//   columns only (kz = 0 and kz = Nz/2), 8 = keep only r_lo < K <= r_hi,
//   16 = keep the ring r_lo < K <= r_hi, a_lo < theta <= a_hi (polar angle
//   to the axis dirn), 32 = keep the cylindrical ring r_lo < K <= r_hi,
//   a_lo < Kpll <= a_hi.
template <typename T>
hipError_t launch_xinv_state(const Grid &g, const Band &b, double *const *V, int ncomp, void *const *U,
                             double *slots, int flags, const void *twx, hipStream_t st, double r_lo = 0.0,
                             double r_hi = 0.0, double a_lo = 0.0, double a_hi = 0.0, int dirn = 3);

// x-forward of T work fields into FP64 spectral outputs, scaled.
template <typename T>
hipError_t launch_xfwd_spec(const Grid &g, void *const *W, int ncomp, double *const *out, double scale,
                            const void *twx, hipStream_t st);

// z-line c2r of T work fields into FP64 real arrays (Nx,Ny,Nz+2).
template <typename T>
hipError_t launch_z_c2r(const Grid &g, void *const *W, int ncomp, double *const *R, const void *twm,
                        const void *twn, hipStream_t st);

// z-line r2c of FP64 real arrays into T work fields.
template <typename T>
hipError_t launch_z_r2c(const Grid &g, const double *const *R, int ncomp, void *const *W, const void *twm,
                        const void *twn, hipStream_t st);

// z pass of the NS nonlinear term: 3 velocity lines in, 6 product lines out
// (order xx, yy, zz, xy, yz, zx).  giver 2: the five trace-shifted products
// xx - zz, yy - zz, xy, yz, zx (for a nonlinear term that is only used
// projected).  giver 1: 6 lines in (u then w), 9 products out (u_j w_i,
// order j-major: (x,x),(x,y),(x,z),(y,x)...); giver 3: the first eight of
// those with P_xx and P_yy less P_zz (the MHD stage, whose velocity term is
// only used projected).
// tape (NS form only): also store the three real velocity lines, each line as
// Nz/2 packed points (x[2n] + i x[2n+1]) at line*(Nz/2).
template <typename T>
hipError_t launch_z_products(const Grid &g, void *const *U, void *const *P, int giver, const void *twm,
                             const void *twn, hipStream_t st, void *const *tape = nullptr);

struct RkParams {
    int stage;            // 1..4; 0 = write the nonlinear term only; -1 = giver
                          // form (9 products u_j w_i), nonlinear term only
    int write_rhs;        // stage 0: apply force/projection/negation (Compute_rhs)
    double dt, fh, f1;    // fh = -nu*0.5*dt, f1 = -nu*1.0*dt
    double fhh, f1h;      // the same with nu_hyper (hyperviscosity)
    int hq;               // hyper_exponent q (0: none)
    double scale;         // 1/(Nx Ny Nz)
    int *nonfinite;       // stage 4: set to 1 if the new state holds a
                          // non-finite value (null: not checked)
    int trace_free;       // the products are the five trace-shifted ones
                          // (stages 1-4, or stage 0 with write_rhs)
};

// Fused x pass: forward x-FFT of the 6 (giver: 9) products, nonlinear term, RHS and the
// Lawson RK4 stage algebra, next stage input, dealias, inverse x-FFT.
// x_tile_columns selects the tile width: the LDS-budget width, or half of it
// (smaller blocks, more of them resident per CU).
template <typename T>
hipError_t launch_x_nlin(const Grid &g, const Band &b, const RkParams &rk, void *const *P, double *const *V,
                         double *const *acc, void *const *U, const double *slots_in, double *slots_out,
                         double *const *out, const void *twx, hipStream_t st, int x_tile_columns = 0);

hipError_t launch_reality(const Grid &g, double *const *V, int ncomp, hipStream_t st);
// n ints at p set to v by a kernel (no memset node, so it may be captured)
hipError_t launch_fill_int(int *p, int n, int v, hipStream_t st);
hipError_t launch_dealias(const Grid &g, double *const *V, int ncomp, hipStream_t st);

}  // namespace tg
