// tg_aniso.h -- anisotropic wavenumber geometry shared by the giver filters
// of the x-inverse pass and the ring / cylindrical-ring reductions.
// dirn is the anisotropy axis (global.field.anisotropy_dirn, 1..3).
#pragma once
#include "tg_kernels.h"

namespace tg {

constexpr double kAnisEps = 1e-15;  // myconstant.MYEPS (Global.cc:62)

// AnisKpll / AnisKperp (FFF_pencil_inline.h:696-725)
__device__ __forceinline__ double aniso_kpll(const Grid &g, int kx, int ky, int kz, int dirn)
{
    return dirn == 1 ? kx * g.kf[0] : (dirn == 2 ? ky * g.kf[1] : kz * g.kf[2]);
}

__device__ __forceinline__ double aniso_kperp(const Grid &g, int kx, int ky, int kz, int dirn)
{
    const double a = kx * g.kf[0], b = ky * g.kf[1], c = kz * g.kf[2];
    return dirn == 1 ? sqrt(b * b + c * c) : (dirn == 2 ? sqrt(a * a + c * c) : sqrt(a * a + b * b));
}

// AnisKvect_polar_angle via Get_polar_angle (FFF_pencil_inline.h:819-827,
// basicfn_inline.h:357-378): angle to the anisotropy axis in [0, pi]
__device__ __forceinline__ double aniso_polar(const Grid &g, int kx, int ky, int kz, int dirn)
{
    const double kpll = aniso_kpll(g, kx, ky, kz, dirn), kperp = aniso_kperp(g, kx, ky, kz, dirn);
    const double pi = 3.14159265358979323846;
    double t;
    if (fabs(kperp) > kAnisEps) t = fabs(kpll) > kAnisEps ? atan(kperp / kpll) : pi / 2;
    else t = kpll >= 0 ? 0.0 : pi;
    return t >= 0 ? t : t + pi;
}

// AnisKvect_azimuthal_angle via Get_azimuthal_angle (FFF_pencil_inline.h:
// 729-757,838-845, basicfn_inline.h:389-407): the angle of (Kh1, Kh2) in
// [0, 2 pi), 0 on the axis
__device__ __forceinline__ double aniso_azimuth(const Grid &g, int kx, int ky, int kz, int dirn)
{
    const double a = kx * g.kf[0], b = ky * g.kf[1], c = kz * g.kf[2];
    const double h1 = dirn == 1 ? b : (dirn == 2 ? c : a), h2 = dirn == 1 ? c : (dirn == 2 ? a : b);
    const double kperp = sqrt(h1 * h1 + h2 * h2);
    if (!(fabs(kperp) > kAnisEps)) return 0.0;
    const double t = acos(h1 / kperp);
    return h2 >= 0 ? t : 2 * 3.14159265358979323846 - t;
}

// first i in [0, n] with arr[i] >= x (n when none)
__device__ __forceinline__ int first_ge(const double *arr, int n, double x)
{
    int lo = 0, hi = n;
    if (!(arr[n] >= x)) return n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (arr[mid] >= x) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Get_sector_index (shell_etc_indices.cc:58-66): angle folded into [0, pi/2]
__device__ __forceinline__ int sector_of(double theta, const double *angles, int nsec)
{
    const double pi = 3.14159265358979323846;
    if (theta > pi / 2) theta = pi - theta;
    return fabs(theta) < kAnisEps ? 1 : first_ge(angles, nsec, theta);
}

// Get_slab_index (shell_etc_indices.cc:93-98)
__device__ __forceinline__ int slab_of(double kpll, const double *kpll_array, int nslabs)
{
    return fabs(kpll) < kAnisEps ? 1 : first_ge(kpll_array, nslabs, fabs(kpll));
}

}  // namespace tg
