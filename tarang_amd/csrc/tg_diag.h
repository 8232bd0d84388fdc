// tg_diag.h -- reduction and initial-field launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "tg_kernels.h"

namespace tg {

enum {
    RED_ENERGY = 0,          // sum w|A|^2, w = 1/2 on kz = 0
    RED_SN = 1,              // sum m K^n |A|^2 (K <= kcut; origin skipped for n = 0)
    RED_SPECTRUM = 2,        // bins ceil(K) <= kcut: [m K^n |A|^2 K^2, 2m]
    RED_SHELL_ALL = 3,       // bins first radii >= K, K <= kcut: 2m Re(A.B*)
    RED_SHELL_ALL_NOCUT = 4, // same without the K cut
    RED_SHELL_SINGLE = 5,    // sum over r_lo < K <= r_hi of 2m Re(A.B*)
    RED_CROSS = 6,           // sum w Re(A B*), w = 1/2 on kz = 0
    // anisotropic bins (RedAniso): shell-major, (sub-bins + 1) per shell
    RED_RING_SPEC = 7,       // [c][ceil(K) <= kcut][sector]: m K^n |A_c|^2, c < ncomp
    RED_CYL_SPEC = 8,        // [q][ceil(Kperp) <= kcut][slab]: m K^n |A_pll|^2, m K^n (|A|^2 - |A_pll|^2)
    RED_RING_MULT = 9,       // [shell of K][sector], 0 < K <= kcut: 2m Re(A.B*)
    RED_CYL_MULT = 10,       // [shell of Kperp][slab], Kperp <= kcut: 2m Re(A.B*)
};

// Bins of the anisotropic modes: sub-bin arrays (sector angles or Kpll
// edges) in aux, the shell radii of the transfer modes in radii_dev, and a
// window [bin_lo, bin_lo + nbins) of the flattened bins (the launcher runs
// one reduction per window that fits the LDS and partial buffers).
struct RedAniso {
    const double *aux;  // sub-bin edges [0..nsub]
    int nsub;           // sectors or slabs
    int nshell;         // shells (radii_dev[0..nshell]) of the transfer modes
    int dirn;           // anisotropy axis 1..3
    int bin_lo;
};

hipError_t launch_reduce(const Grid &g, int mode, const double *const *A, const double *const *B, int ncomp, int n,
                         const double *radii_dev, int nbins, int nq, int kcut, double r_lo, double r_hi,
                         double *partial, int nblocks, double *out_dev, hipStream_t st,
                         const RedAniso *an = nullptr);

hipError_t launch_init_field(const Grid &g, double *const *V, int kind, double k0, double kmin, double kmax,
                             uint64_t seed, hipStream_t st);
hipError_t launch_scale_fields(const Grid &g, double *const *V, const double *target_and_e, hipStream_t st);
// out[0] = max |R| over the Nx*Ny rows of Nz samples (row stride Nz + 2) of a
// real field; a NaN sample gives NaN.  out: 8 bytes of device memory.
hipError_t launch_max_abs_real(const Grid &g, const double *R, double *out, hipStream_t st);

}  // namespace tg
