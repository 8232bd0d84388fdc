// tg_diag.hip -- per-mode reductions (energy, Sn, shell spectrum, shell
// products), the seeded initial field, and the energy-transfer giver pass.
//
// Reductions are two-stage, compensated and deterministic: each block owns a
// fixed slice of modes and reduces it into per-wavefront (sum, compensation)
// bins in LDS (no atomics: lanes that share a bin are summed in a fixed-order
// tree and added with Neumaier compensation), merges the wavefront bins in a
// fixed order and writes one partial per bin; a second kernel sums the
// partials of each bin in block order with Neumaier compensation.
#include <algorithm>

#include "tg_device.h"
#include "tg_kernels.h"

#include "tg_diag.h"
#include "tg_aniso.h"
#include "tg_rand.h"

namespace tg {

__device__ __forceinline__ int kw(int l, int n) { return l <= n / 2 ? l : l - n; }

// Multiplicity_factor (FFF_pencil_inline.h:583-602)
__device__ __forceinline__ double multiplicity(const Grid &g, int lx, int ly, int lz)
{
    double f = lz > 0 ? 1.0 : 0.5;
    const int kx = kw(lx, g.nx), ky = kw(ly, g.ny);
    if (kx == g.nx / 2 || (ky == g.ny / 2 && g.ny > 1)) return 2 * f;
    return f;
}

__device__ __forceinline__ double kmag_l(const Grid &g, int lx, int ly, int lz)
{
    const double a = kw(lx, g.nx) * g.kf[0], b = kw(ly, g.ny) * g.kf[1], c = lz * g.kf[2];
    return sqrt(a * a + b * b + c * c);
}

__device__ __forceinline__ void neumaier(double &s, double &c, double x)
{
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

// first index i with radii[i] >= K (Get_shell_index, shell_etc_indices.cc:49-52)
__device__ __forceinline__ int shell_of(double K, const double *radii, int n)
{
    int lo = 0, hi = n - 1;
    if (!(radii[hi] >= K)) return n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (radii[mid] >= K) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

struct RedArgs {
    const double *A[3];
    const double *B[3];
    const double *radii;
    double *partial;   // [nblocks][nq][nbins]
    Grid g;
    int mode, ncomp, n, nbins, nq, kcut;  // kcut: Max_radius_inside (mode SHELL_ALL) / Min_radius_outside
    double r_lo, r_hi;
    RedAniso an;
};

constexpr int kRedThreads = 256;
constexpr int kRedMaxBins = 1100;

// Sum over the 64 lanes of a wavefront (xor butterfly: every lane ends with
// the same, order-fixed result).
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Adds each active lane's values val[0..nq) into this wavefront's bins
// (sum, compensation) pairs at bin `bin`.  The lanes that share a bin are
// summed by a fixed-order tree, then one lane adds that sum to the bin with
// Neumaier compensation; the loop runs once per distinct bin among the lanes
// (consecutive modes fall in few shells).  Deterministic, no atomics.
template <int NQ>
__device__ __forceinline__ void wave_bin_add(double *bs, double *bc, int nb, int bin, const double (&val)[NQ],
                                             bool active)
{
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int b = __shfl(bin, leader, 64);
        const bool mine = active && bin == b;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const double t = wave_sum(mine ? val[q] : 0.0);
            if (lane == leader) neumaier(bs[q * nb + b], bc[q * nb + b], t);
        }
        pending &= ~__ballot(mine);
    }
}

// wave_bin_add with a run-time number of quantities (nq <= 3)
__device__ __forceinline__ void wave_bin_add_n(double *bs, double *bc, int nb, int nq, int bin, const double *val,
                                               bool active)
{
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int b = __shfl(bin, leader, 64);
        const bool mine = active && bin == b;
        for (int q = 0; q < nq; q++) {
            const double t = wave_sum(mine ? val[q] : 0.0);
            if (lane == leader) neumaier(bs[q * nb + b], bc[q * nb + b], t);
        }
        pending &= ~__ballot(mine);
    }
}

// Modes of block b: a fixed contiguous slice of the flattened mode index.
// Scalar modes: per-lane Neumaier sums, then a fixed-order block tree.
// Binned modes: per-wavefront compensated bins (LDS: [NW][nq][nbins] sums,
// then as many compensations), merged over the wavefronts in order.
template <int MODE>
__global__ void __launch_bounds__(kRedThreads) k_reduce(RedArgs a)
{
    const int NW = blockDim.x / 64;
    extern __shared__ double dyn[];
    const Grid &g = a.g;
    const int wave = threadIdx.x / 64;
    const int nb = a.nbins, nq = a.nq;
    constexpr bool ANIS = MODE == RED_RING_SPEC || MODE == RED_CYL_SPEC || MODE == RED_RING_MULT ||
                          MODE == RED_CYL_MULT;
    constexpr bool BINNED = MODE == RED_SPECTRUM || MODE == RED_SHELL_ALL || MODE == RED_SHELL_ALL_NOCUT || ANIS;
    double *bsum = dyn + (size_t)wave * nq * nb;
    double *bcmp = dyn + (size_t)(NW + wave) * nq * nb;
    if (BINNED) {
        for (int i = threadIdx.x; i < 2 * NW * nq * nb; i += blockDim.x) dyn[i] = 0.0;
        __syncthreads();
    }
    const size_t total = (size_t)g.nx * g.ny * g.nzc;
    const size_t chunk = (total + gridDim.x - 1) / gridDim.x;
    const size_t beg = (size_t)blockIdx.x * chunk;
    const size_t end = beg + chunk < total ? beg + chunk : total;
    double s0 = 0.0, c0 = 0.0;  // single-bin compensated accumulators
    // every lane runs the same number of iterations (the binned modes vote
    // across the wavefront)
    for (size_t i0 = beg; i0 < end; i0 += blockDim.x) {
        const size_t i = i0 + threadIdx.x;
        const bool valid = i < end;
        const size_t ii = valid ? i : beg;
        const int lz = ii % g.nzc;
        const int ly = (ii / g.nzc) % g.ny;
        const int lx = ii / ((size_t)g.nzc * g.ny);
        // FFF: w = 1/2 on kz = 0 (FFF_pencil_energy.cc:59-68); SFF: 2 sum - kz=0
        // plane - kx=0 plane + half the kx=0,kz=0 line (SFF_pencil_energy.cc:
        // 88-104), i.e. w = (lx ? 2 : 1)(lz ? 2 : 1)/2
        const double w = g.basis ? 0.5 * (lx ? 2.0 : 1.0) * (lz ? 2.0 : 1.0) : (lz == 0 ? 0.5 : 1.0);
        if (MODE == RED_ENERGY) {
            const double *A = a.A[0];
            const double re = A[2 * ii], im = A[2 * ii + 1];
            if (valid) neumaier(s0, c0, w * (re * re + im * im));
        } else if (MODE == RED_CROSS) {
            const double *A = a.A[0], *B = a.B[0];
            if (valid) neumaier(s0, c0, w * (A[2 * ii] * B[2 * ii] + A[2 * ii + 1] * B[2 * ii + 1]));
        } else if (MODE == RED_SN) {
            const double K = kmag_l(g, lx, ly, lz);
            if (valid && !(ii == 0 && a.n == 0) && K <= a.kcut) {
                const double *A = a.A[0];
                const double re = A[2 * ii], im = A[2 * ii + 1];
                double kn = 1.0;
                for (int t = 0; t < a.n; t++) kn *= K;
                neumaier(s0, c0, multiplicity(g, lx, ly, lz) * kn * (re * re + im * im));
            }
        } else if (MODE == RED_SPECTRUM) {
            const double K = kmag_l(g, lx, ly, lz);
            const int idx = (int)ceil(K);
            const bool act = valid && idx <= a.kcut;
            const double *A = a.A[0];
            const double re = A[2 * ii], im = A[2 * ii + 1];
            const double m = multiplicity(g, lx, ly, lz);
            double kn = 1.0;
            for (int t = 0; t < a.n; t++) kn *= K;
            const double val[2] = {m * kn * (re * re + im * im) * (K * K), 2 * m};
            wave_bin_add<2>(bsum, bcmp, nb, act ? idx : 0, val, act);
        } else if (MODE == RED_SHELL_ALL || MODE == RED_SHELL_ALL_NOCUT) {
            const double K = kmag_l(g, lx, ly, lz);
            const bool act = valid && (MODE == RED_SHELL_ALL_NOCUT || K <= a.kcut);
            const int sidx = shell_of(K, a.radii, nb);
            double d = 0.0;
            for (int c = 0; c < a.ncomp; c++) {
                const double *A = a.A[c], *B = a.B[c];
                d += A[2 * ii] * B[2 * ii] + A[2 * ii + 1] * B[2 * ii + 1];
            }
            const double val[1] = {2 * multiplicity(g, lx, ly, lz) * d};
            wave_bin_add<1>(bsum, bcmp, nb, sidx, val, act);
        } else if (ANIS) {
            const RedAniso &an = a.an;
            const int kx = kw(lx, g.nx), ky = kw(ly, g.ny);
            const double K = kmag_l(g, lx, ly, lz);
            const double m = multiplicity(g, lx, ly, lz);
            double val[3] = {0.0, 0.0, 0.0};
            int shell = 0, sub = 0;
            bool act = valid;
            if (MODE == RED_RING_SPEC || MODE == RED_RING_MULT) {
                // Compute_local_ring_spectrum (universal_energy.cc:552-590) and
                // Local_ring_mult_all (universal_ET.cc:205-230)
                if (MODE == RED_RING_SPEC) {
                    shell = (int)ceil(K);
                    act = act && K > kAnisEps && shell <= a.kcut;
                } else {
                    act = act && K <= a.kcut && K > kAnisEps;
                    shell = first_ge(a.radii, an.nshell, K);
                }
                sub = sector_of(aniso_polar(g, kx, ky, lz, an.dirn), an.aux, an.nsub);
            } else {
                // Compute_local_cylindrical_ring_spectrum (universal_energy.cc:
                // 868-906) and Local_cyl_ring_mult_all (universal_ET.cc:335-362)
                const double kperp = aniso_kperp(g, kx, ky, lz, an.dirn);
                if (MODE == RED_CYL_SPEC) {
                    shell = (int)ceil(kperp);
                    act = act && shell <= a.kcut;
                } else {
                    act = act && kperp <= a.kcut;
                    shell = first_ge(a.radii, an.nshell, kperp);
                }
                sub = slab_of(aniso_kpll(g, kx, ky, lz, an.dirn), an.aux, an.nsub);
            }
            if (MODE == RED_RING_SPEC || MODE == RED_CYL_SPEC) {
                double kn = 1.0;
                for (int t = 0; t < a.n; t++) kn *= K;
                const double f = m * kn;
                if (MODE == RED_RING_SPEC) {
                    for (int c = 0; c < a.nq; c++) {
                        const double *A = a.A[c];
                        val[c] = f * (A[2 * ii] * A[2 * ii] + A[2 * ii + 1] * A[2 * ii + 1]);
                    }
                } else {
                    double v2 = 0.0, p2 = 0.0;
                    for (int c = 0; c < 3; c++) {
                        const double *A = a.A[c];
                        const double s2 = A[2 * ii] * A[2 * ii] + A[2 * ii + 1] * A[2 * ii + 1];
                        v2 += s2;
                        if (c == an.dirn - 1) p2 = s2;
                    }
                    val[0] = f * p2;
                    val[1] = f * (v2 - p2);
                }
            } else {
                double d = 0.0;
                for (int c = 0; c < 3; c++) {
                    const double *A = a.A[c], *B = a.B[c];
                    d += A[2 * ii] * B[2 * ii] + A[2 * ii + 1] * B[2 * ii + 1];
                }
                val[0] = 2 * m * d;
            }
            const int bin = shell * (an.nsub + 1) + sub - an.bin_lo;
            act = act && bin >= 0 && bin < nb;
            wave_bin_add_n(bsum, bcmp, nb, nq, act ? bin : 0, val, act);
        } else if (MODE == RED_SHELL_SINGLE) {
            const double K = kmag_l(g, lx, ly, lz);
            if (valid && K > a.r_lo && K <= a.r_hi) {
                double d = 0.0;
                for (int c = 0; c < a.ncomp; c++) {
                    const double *A = a.A[c], *B = a.B[c];
                    d += A[2 * ii] * B[2 * ii] + A[2 * ii + 1] * B[2 * ii + 1];
                }
                neumaier(s0, c0, 2 * multiplicity(g, lx, ly, lz) * d);
            }
        }
    }
    if (!BINNED) {
        // block tree reduction of (s0 + c0)
        double v = s0 + c0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0) dyn[wave] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < NW; w++) t += dyn[w];
            a.partial[blockIdx.x] = t;
        }
        return;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nq * nb; i += blockDim.x) {
        double t = dyn[i], c = dyn[(size_t)NW * nq * nb + i];
        for (int w = 1; w < NW; w++) {
            neumaier(t, c, dyn[(size_t)w * nq * nb + i]);
            c += dyn[(size_t)(NW + w) * nq * nb + i];
        }
        a.partial[(size_t)blockIdx.x * nq * nb + i] = t + c;
    }
}

// out[i] = sum over blocks of partial[b][i], compensated, block order.
__global__ void k_reduce_final(const double *partial, int nblocks, int width, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= width) return;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < nblocks; b++) neumaier(s, c, partial[(size_t)b * width + i]);
    out[i] = s + c;
}

hipError_t launch_reduce(const Grid &g, int mode, const double *const *A, const double *const *B, int ncomp,
                         int n, const double *radii_dev, int nbins, int nq, int kcut, double r_lo, double r_hi,
                         double *partial, int nblocks, double *out_dev, hipStream_t st, const RedAniso *an)
{
    RedArgs a{};
    if (an) a.an = *an;
    for (int c = 0; c < ncomp; c++) {
        a.A[c] = A[c];
        a.B[c] = B ? B[c] : nullptr;
    }
    a.radii = radii_dev;
    a.partial = partial;
    a.g = g;
    a.mode = mode;
    a.ncomp = ncomp;
    a.n = n;
    a.nbins = nbins;
    a.nq = nq;
    a.kcut = kcut;
    a.r_lo = r_lo;
    a.r_hi = r_hi;
    const bool single = mode == RED_ENERGY || mode == RED_SN || mode == RED_SHELL_SINGLE || mode == RED_CROSS;
    const int width = single ? 1 : nq * nbins;
    if (!single && nbins > kRedMaxBins) return hipErrorInvalidValue;
    // binned: (sum, compensation) per bin and wavefront, within 64 KB of LDS
    int threads = kRedThreads;
    if (!single) {
        const size_t per_wave = (size_t)2 * nq * nbins * sizeof(double);
        int nw = (int)std::min<size_t>(kRedThreads / 64, 65536 / per_wave);
        threads = 64 * (nw < 1 ? 1 : nw);
    }
    const size_t shm = single ? 64 : (size_t)(threads / 64) * 2 * nq * nbins * sizeof(double);
    switch (mode) {
        case RED_ENERGY: hipLaunchKernelGGL(k_reduce<RED_ENERGY>, dim3(nblocks), dim3(threads), shm, st, a); break;
        case RED_SN: hipLaunchKernelGGL(k_reduce<RED_SN>, dim3(nblocks), dim3(threads), shm, st, a); break;
        case RED_SPECTRUM: hipLaunchKernelGGL(k_reduce<RED_SPECTRUM>, dim3(nblocks), dim3(threads), shm, st, a); break;
        case RED_SHELL_ALL: hipLaunchKernelGGL(k_reduce<RED_SHELL_ALL>, dim3(nblocks), dim3(threads), shm, st, a); break;
        case RED_SHELL_ALL_NOCUT:
            hipLaunchKernelGGL(k_reduce<RED_SHELL_ALL_NOCUT>, dim3(nblocks), dim3(threads), shm, st, a);
            break;
        case RED_SHELL_SINGLE:
            hipLaunchKernelGGL(k_reduce<RED_SHELL_SINGLE>, dim3(nblocks), dim3(threads), shm, st, a);
            break;
        case RED_CROSS: hipLaunchKernelGGL(k_reduce<RED_CROSS>, dim3(nblocks), dim3(threads), shm, st, a); break;
        case RED_RING_SPEC:
            hipLaunchKernelGGL(k_reduce<RED_RING_SPEC>, dim3(nblocks), dim3(threads), shm, st, a);
            break;
        case RED_CYL_SPEC: hipLaunchKernelGGL(k_reduce<RED_CYL_SPEC>, dim3(nblocks), dim3(threads), shm, st, a); break;
        case RED_RING_MULT:
            hipLaunchKernelGGL(k_reduce<RED_RING_MULT>, dim3(nblocks), dim3(threads), shm, st, a);
            break;
        case RED_CYL_MULT: hipLaunchKernelGGL(k_reduce<RED_CYL_MULT>, dim3(nblocks), dim3(threads), shm, st, a); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3((width + 127) / 128), dim3(128), 0, st, partial, nblocks, width, out_dev);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Seeded initial field (same definition as the oracle's or_init_field).
// ---------------------------------------------------------------------------

struct InitArgs {
    double *V[3];
    Grid g;
    int kind;
    double k0, kmin, kmax;
    uint64_t seed;
};

__global__ void k_init_field(InitArgs a)
{
    const Grid &g = a.g;
    const size_t total = (size_t)g.nx * g.ny * g.nzc;
    const double twopi = 6.283185307179586476925286766559;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int lz = i % g.nzc;
        const int ly = (i / g.nzc) % g.ny;
        const int lx = i / ((size_t)g.nzc * g.ny);
        const int kx = kw(lx, g.nx), ky = kw(ly, g.ny);
        const double Kx = kx * g.kf[0], Ky = ky * g.kf[1], Kz = lz * g.kf[2];
        const double K = sqrt(Kx * Kx + Ky * Ky + Kz * Kz);
        double Ek = 0.0;
        if (K > 0.0) {
            if (a.kind == 0) Ek = pow(K, 4.0) * exp(-2.0 * (K / a.k0) * (K / a.k0));
            else if (K >= a.kmin && K <= a.kmax) Ek = pow(K, -5.0 / 3.0);
        }
        if (kx == g.nx / 2 || ky == g.ny / 2 || lz == g.nz / 2) Ek = 0.0;
        double v[6] = {0, 0, 0, 0, 0, 0};
        if (Ek != 0.0) {
            const double amp = sqrt(Ek / (2.0 * 3.14159265358979323846 * K * K));
            const double kp = sqrt(Kx * Kx + Ky * Ky);
            double e1[3], e2[3];
            if (kp > 0.0) { e1[0] = Ky / kp; e1[1] = -Kx / kp; e1[2] = 0.0; }
            else { e1[0] = 1.0; e1[1] = 0.0; e1[2] = 0.0; }
            e2[0] = (Ky * e1[2] - Kz * e1[1]) / K;
            e2[1] = (Kz * e1[0] - Kx * e1[2]) / K;
            e2[2] = (Kx * e1[1] - Ky * e1[0]) / K;
            const uint64_t base = (uint64_t)i * 4u;
            double a1r, a1i, a2r, a2i;
            if (a.kind == 0) {
                const double u1 = 1.0 - uniform01(a.seed, base + 0), u2 = uniform01(a.seed, base + 1);
                const double u3 = 1.0 - uniform01(a.seed, base + 2), u4 = uniform01(a.seed, base + 3);
                const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
                a1r = r1 * cos(twopi * u2) / 2.0;
                a1i = r1 * sin(twopi * u2) / 2.0;
                a2r = r2 * cos(twopi * u4) / 2.0;
                a2i = r2 * sin(twopi * u4) / 2.0;
            } else {
                const double p1 = uniform01(a.seed, base + 0), p2 = uniform01(a.seed, base + 1);
                const double r = sqrt(2.0);
                a1r = cos(twopi * p1) / r;
                a1i = sin(twopi * p1) / r;
                a2r = cos(twopi * p2) / r;
                a2i = sin(twopi * p2) / r;
            }
            for (int c = 0; c < 3; c++) {
                v[2 * c] = amp * (a1r * e1[c] + a2r * e2[c]);
                v[2 * c + 1] = amp * (a1i * e1[c] + a2i * e2[c]);
            }
        }
        for (int c = 0; c < 3; c++) {
            a.V[c][2 * i] = v[2 * c];
            a.V[c][2 * i + 1] = v[2 * c + 1];
        }
    }
}

__global__ void k_scale_fields(double *V0, double *V1, double *V2, size_t n, const double *target_and_e)
{
    // target_and_e[0] = target energy, [1..3] = component energies
    const double e = target_and_e[1] + target_and_e[2] + target_and_e[3];
    const double s = e > 0.0 ? sqrt(target_and_e[0] / e) : 1.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        V0[i] *= s;
        V1[i] *= s;
        V2[i] *= s;
    }
}

hipError_t launch_init_field(const Grid &g, double *const *V, int kind, double k0, double kmin, double kmax,
                             uint64_t seed, hipStream_t st)
{
    InitArgs a{};
    for (int c = 0; c < 3; c++) a.V[c] = V[c];
    a.g = g;
    a.kind = kind;
    a.k0 = k0;
    a.kmin = kmin;
    a.kmax = kmax;
    a.seed = seed;
    hipLaunchKernelGGL(k_init_field, dim3(4096), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_scale_fields(const Grid &g, double *const *V, const double *target_and_e, hipStream_t st)
{
    const size_t n = (size_t)g.nx * g.ny * g.nzc * 2;
    hipLaunchKernelGGL(k_scale_fields, dim3(4096), dim3(256), 0, st, V[0], V[1], V[2], n, target_and_e);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Maximum of |u| over a real field (Get_local_max_real_space, FluidVF.cc:
// 499-509).  The maximum of non-negative doubles is the maximum of their bit
// patterns as unsigned integers (NaN above every number), so each wave
// reduces with shuffles and adds its maximum with one integer atomic max: the
// result does not depend on the order.
// ---------------------------------------------------------------------------
__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

__global__ void k_max_abs_real(const double *R, size_t nrows, int nz, int stride, unsigned long long *out)
{
    double m = 0.0;
    const size_t n = nrows * (size_t)nz;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = max_nan(m, fabs(R[(i / nz) * stride + i % nz]));
    for (int o = 32; o > 0; o >>= 1) m = max_nan(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

hipError_t launch_max_abs_real(const Grid &g, const double *R, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, (unsigned long long *)out);
    const size_t nrows = (size_t)g.nx * g.ny;
    const int blocks = (int)std::min<size_t>(4096, (nrows * g.nz + 255) / 256);
    hipLaunchKernelGGL(k_max_abs_real, dim3(blocks), dim3(256), 0, st, R, nrows, g.nz, g.nz + 2,
                       (unsigned long long *)out);
    return hipGetLastError();
}

}  // namespace tg
