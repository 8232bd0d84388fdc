// tg_kernels.hip -- gfx950 kernels of the pseudo-spectral FFF hot path.
//
// HBM layouts (all complex fields interleaved re/im):
//   spectral state  V(lx,ly,lz)        (Nx, Ny, Nz/2+1) FP64, kz fastest
//   work fields     W(x | ly | lz ...) (Nx, Ny, Nz/2+1) in the transform
//                   precision T; the same shape serves every intermediate
//                   of the 3D transform (x, y and z passes work in place on
//                   it, each pass finishing one axis).
//   real fields     R(x,y,z)           (Nx, Ny, Nz+2) FP64, z fastest.
//
// Passes:
//   x pass  tiles of all Nx rows x CB contiguous (ly,lz) columns, staged
//           through LDS (coalesced CB-wide row segments), one line per
//           TPL = Nx/8 threads.
//   y pass  tiles of all Ny rows x CB contiguous lz columns of one lx.
//   z pass  whole z-lines (contiguous), TPL = Nz/16 threads per line; a
//           real line of Nz points is packed as Nz/2 complex points.
#include <algorithm>
#include <type_traits>

#include "tg_device.h"
#include "tg_kernels.h"
#include "tg_pass.h"
#include "tg_aniso.h"
#include "tg_rand.h"

namespace tg {

// ---------------------------------------------------------------------------
// Column FFT along a strided axis (the y pass, or a plain x pass).
// ---------------------------------------------------------------------------
struct ColArgs {
    void *f[9];
    const void *tw;
    long long outer_stride;  // elements between tiles of different outer index
    long long row_stride;    // elements between consecutive rows (transform axis)
    int ncol;                // contiguous columns per row
    int nouter;
};

template <typename T, int L, int DIR, int CB>
__global__ void __launch_bounds__(CB *(L / 8)) k_colfft(ColArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    __shared__ cpx<T> lds[CB * LP];
    // xcd_tile: the column blocks of one outer index run in turn on one XCD,
    // so the rows' 128-byte lines that two blocks share (CB columns of fewer
    // than 128 bytes) are fetched into one L2
    const int nblk = (a.ncol + CB - 1) / CB;
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    if (b >= a.nouter * nblk) return;
    const int outer = b / nblk;
    const int c0 = (b % nblk) * CB;
    cpx<T> *f = (cpx<T> *)a.f[blockIdx.y] + (size_t)outer * a.outer_stride;
    const cpx<T> *tw = (const cpx<T> *)a.tw;
    // thread t moves column t % CB, rows t / CB + TPL i: the eight loads are
    // issued before the first LDS write (a column past the last one loads
    // the last one, stages zeros and is not stored)
    const int sc = threadIdx.x % CB, sr = threadIdx.x / CB, scol = c0 + sc;
    const bool live = scol < a.ncol;
    {
        const cpx<T> *p = f + (size_t)sr * a.row_stride + min(scol, a.ncol - 1);
        const size_t step = (size_t)TPL * a.row_stride;
        cpx<T> x[8];
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = mk<T>(p[i * step].x, p[i * step].y);
#pragma unroll
        for (int i = 0; i < 8; i++) lds[sc * LP + lpad(sr + TPL * i)] = live ? x[i] : mk<T>(0, 0);
    }
    lds_barrier();
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    cpx<T> v[8];
#pragma unroll
    for (int s = 0; s < 8; s++) v[s] = lds[c * LP + lpad(j + TPL * s)];
    lds_barrier();
    fft_line<L, DIR>(v, lds + c * LP, tw, j);
#pragma unroll
    for (int s = 0; s < 8; s++) lds[c * LP + lpad(j + TPL * s)] = v[s];
    lds_barrier();
    if (live) {
        cpx<T> *p = f + (size_t)sr * a.row_stride + scol;
        const size_t step = (size_t)TPL * a.row_stride;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const cpx<T> v = lds[sc * LP + lpad(sr + TPL * i)];
            p[i * step] = mk<T>(v.x, v.y);
        }
    }
}

// ---------------------------------------------------------------------------
// x-inverse of FP64 spectral fields into T work fields (first pass of the
// 3D inverse transform; stage-1 input of an RK4 step).
// ---------------------------------------------------------------------------
struct XinvArgs {
    double *V[3];
    void *U[3];
    double *slots;
    const void *tw;
    Grid g;
    Band b;
    int ncomp, flags;
    double r_lo, r_hi;  // flags & 8: keep only r_lo < K <= r_hi (Fill_array_shell)
    // flags & 16: ring r_lo < K <= r_hi, a_lo < theta <= a_hi (Fill_array_ring);
    // flags & 32: cylindrical ring r_lo < K <= r_hi, a_lo < Kpll <= a_hi
    // (Fill_array_cylindrical_ring); theta, Kpll about the axis dirn
    double a_lo, a_hi;
    int dirn;
};

template <typename T, int L, int CB>
__global__ void __launch_bounds__(CB *(L / 8)) k_xinv_state(XinvArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    __shared__ __attribute__((aligned(16))) char raw[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    cpx<double> *ldd = (cpx<double> *)raw;
    cpx<T> *ldt = (cpx<T> *)raw;
    const Grid &g = a.g;
    const bool patch = (a.flags & 4) != 0;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= x_tiles(g, CB, patch)) return;
    tile_columns<CB>(colw, cols, g, patch, tile);
    lds_barrier();
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int cw = colw[c];
    const int ly = cw >= 0 ? cw / g.nzp : 0, lz = cw >= 0 ? cw % g.nzp : 0;
    const int ky = kwave(ly, g.ny);
    const cpx<T> *tw = (const cpx<T> *)a.tw;
    const int ncw = g.ny * g.nzp, ncs = g.ny * g.nzc;
    for (int comp = 0; comp < a.ncomp; comp++) {
        stage_in<L, CB>(ldd, (const cpx<double> *)a.V[comp], cols, ncs);
        lds_barrier();
        cpx<T> v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int lx = j + TPL * s;
            cpx<double> d = ldd[c * LP + lpad(lx)];
            if (cw >= 0 && (a.flags & 2)) {
                const int kx = kwave(lx, g.nx);
                const int sl = band_slot(a.b, kx, ky, lz);
                if (sl >= 0) {
                    const double K = kmag(g, kx, ky, lz);
                    if (K > a.b.r_in && K <= a.b.r_out) ((cpx<double> *)a.slots)[sl * 3 + comp] = d;
                }
            }
            if ((a.flags & 1) && dealias_zero(g, lx, ly, lz)) d = mk<double>(0, 0);
            if (a.flags & 8) {
                const double K = kmag(g, kwave(lx, g.nx), ky, lz);
                if (!(K > a.r_lo && K <= a.r_hi)) d = mk<double>(0, 0);
            }
            if (a.flags & 48) {
                const int kx = kwave(lx, g.nx);
                const double K = kmag(g, kx, ky, lz);
                // universal_ET.cc:909-911 (ring) and :929-931 (cylindrical ring)
                const double q = (a.flags & 16) ? aniso_polar(g, kx, ky, lz, a.dirn) : aniso_kpll(g, kx, ky, lz, a.dirn);
                const bool in = K > a.r_lo && K <= a.r_hi &&
                                ((q > a.a_lo && q <= a.a_hi) || (fabs(q) < kAnisEps && a.a_lo < kAnisEps));
                if (!in) d = mk<double>(0, 0);
            }
            v[s] = cvt<T>(d);
        }
        lds_barrier();
        fft_line<L, +1>(v, ldt + c * LP, tw, j);
#pragma unroll
        for (int s = 0; s < 8; s++) ldt[c * LP + lpad(j + TPL * s)] = v[s];
        lds_barrier();
        stage_out<L, CB>(ldt, (cpx<T> *)a.U[comp], colw, ncw);
        lds_barrier();
    }
}

// ---------------------------------------------------------------------------
// x-forward of T work fields into FP64 spectral outputs (last pass of the 3D
// forward transform), scaled.
// ---------------------------------------------------------------------------
struct XfwdArgs {
    void *W[3];
    double *out[3];
    const void *tw;
    Grid g;
    double scale;
    int ncomp;
};

template <typename T, int L, int CB>
__global__ void __launch_bounds__(CB *(L / 8)) k_xfwd_spec(XfwdArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    __shared__ __attribute__((aligned(16))) char raw[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    cpx<double> *ldd = (cpx<double> *)raw;
    cpx<T> *ldt = (cpx<T> *)raw;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= x_tiles(a.g, CB, false)) return;
    tile_columns<CB>(colw, cols, a.g, false, tile);
    lds_barrier();
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const cpx<T> *tw = (const cpx<T> *)a.tw;
    const int ncw = a.g.ny * a.g.nzp, ncs = a.g.ny * a.g.nzc;
    for (int comp = 0; comp < a.ncomp; comp++) {
        stage_in<L, CB>(ldt, (const cpx<T> *)a.W[comp], colw, ncw);
        lds_barrier();
        cpx<T> v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) v[s] = ldt[c * LP + lpad(j + TPL * s)];
        lds_barrier();
        fft_line<L, -1>(v, ldt + c * LP, tw, j);
        lds_barrier();
#pragma unroll
        for (int s = 0; s < 8; s++) ldd[c * LP + lpad(j + TPL * s)] = a.scale * cvt<double>(v[s]);
        lds_barrier();
        stage_out<L, CB>(ldd, (cpx<double> *)a.out[comp], cols, ncs);
        lds_barrier();
    }
}

template <typename T, int M>
__global__ void __launch_bounds__(256) k_z_c2r(ZArgs a)
{
    constexpr int TPL = M / 8;
    constexpr int LPB = z_lines_per_block<M>();
    constexpr int LP = lpad_len(M);
    __shared__ cpx<T> lds[LPB * LP];
    const int li = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const size_t line = (size_t)blockIdx.x * LPB + li;
    const size_t nlines = (size_t)a.g.nx * a.g.ny;
    const size_t lsafe = line < nlines ? line : nlines - 1;
    for (int comp = 0; comp < a.ncomp; comp++) {
        const cpx<T> *X = (const cpx<T> *)a.in[comp] + lsafe * a.g.nzp;
        cpx<T> v[8];
        zline_c2r<T, M>(v, X, lds + li * LP, (const cpx<T> *)a.twm, (const cpx<T> *)a.twn, j);
        if (line < nlines) {
            double *R = (double *)a.out[comp] + line * (size_t)(2 * M + 2);
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int n = j + TPL * s;
                ((cpx<double> *)R)[n] = cvt<double>(v[s]);
            }
            if (j == 0) ((cpx<double> *)R)[M] = mk<double>(0, 0);
        }
        line_sync<M / 8>();
    }
}

template <typename T, int M>
__global__ void __launch_bounds__(256) k_z_r2c(ZArgs a)
{
    constexpr int TPL = M / 8;
    constexpr int LPB = z_lines_per_block<M>();
    constexpr int LP = lpad_len(M);
    __shared__ cpx<T> lds[LPB * LP];
    const int li = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const size_t line = (size_t)blockIdx.x * LPB + li;
    const size_t nlines = (size_t)a.g.nx * a.g.ny;
    const size_t lsafe = line < nlines ? line : nlines - 1;
    for (int comp = 0; comp < a.ncomp; comp++) {
        const cpx<double> *R = (const cpx<double> *)((const double *)a.in[comp] + lsafe * (size_t)(2 * M + 2));
        cpx<T> v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) v[s] = cvt<T>(R[j + TPL * s]);
        cpx<T> *X = (cpx<T> *)a.out[comp] + lsafe * a.g.nzp;
        zline_r2c_post<T, M>(v, X, lds + li * LP, (const cpx<T> *)a.twm, (const cpx<T> *)a.twn, j, line < nlines);
    }
}

// z pass of the nonlinear term.  NS: products u_a u_b for the pairs
// (0,0),(1,1),(2,2),(0,1),(1,2),(2,0) -- the six real-space products of
// Compute_nlin_diag / Compute_nlin_offdiag (compute_diag_offdiag.cc:47-112);
// a.ncomp == 5: the trace-shifted five u_x^2 - u_z^2, u_y^2 - u_z^2, u_x u_y,
// u_y u_z, u_z u_x (the nonlinear term less the gradient i K F[u_z^2], which
// the pressure projection removes).
// Giver: out[3*i + j] = u_j w_i (compute_diag_offdiag.cc:329-414).
// FP32 lines of 256 points (512-point z axis): four waves per SIMD (the
// allocation fits 128 VGPRs without spills and the pass runs 6% faster at
// 512^3 than with the three waves the unconstrained allocation gives).
// FIVE: the NS form with the five trace-shifted products (an instance of its
// own: with both NS forms in one kernel the FP32 instance spilled under its
// four-wave bound); with GIVER, the eight trace-shifted products of the MHD
// stage: P_xx - P_zz, P_yy - P_zz and the six off-diagonal ones (P_zz is
// not stored).
template <typename T, int M, bool GIVER, bool FIVE = false>
__global__ void __launch_bounds__(256, (sizeof(T) == 4 && !GIVER && M == 256) ? 4 : 1) k_z_products(ZArgs a)
{
    constexpr int TPL = M / 8;
    constexpr int LPB = z_lines_per_block<M>();
    constexpr int LP = lpad_len(M);
    __shared__ cpx<T> lds[LPB * LP];
    const int li = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const size_t line = (size_t)blockIdx.x * LPB + li;
    const size_t nlines = (size_t)a.g.nx * a.g.ny;
    const bool valid = line < nlines;
    const size_t lsafe = valid ? line : nlines - 1;
    const cpx<T> *twm = (const cpx<T> *)a.twm, *twn = (const cpx<T> *)a.twn;
    cpx<T> *L = lds + li * LP;
    // the three input lines are loaded before the first transform: their
    // loads are in flight together
    cpx<T> u[3][8], xm[3];
#pragma unroll
    for (int c = 0; c < 3; c++) zline_load<T, M>(u[c], xm[c], (const cpx<T> *)a.in[c] + lsafe * a.g.nzp, j);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        zline_c2r_from<T, M>(u[c], xm[c], L, twm, twn, j);
        line_sync<M / 8>();
    }
    if (!GIVER && a.tape[0] && valid) {
        // the adjoint's record of u = I(D V) for this stage
#pragma unroll
        for (int c = 0; c < 3; c++) {
            cpx<T> *tp = (cpx<T> *)a.tape[c] + line * M;
#pragma unroll
            for (int s = 0; s < 8; s++) tp[j + TPL * s] = u[c][s];
        }
    }
    auto emit = [&](cpx<T> *p, int o) {
        cpx<T> *X = (cpx<T> *)a.out[o] + lsafe * a.g.nzp;
        zline_r2c_post<T, M>(p, X, L, twm, twn, j, valid);
    };
    if constexpr (!GIVER && FIVE) {
        // the cross products first, then the two shifted squares (u_z is
        // last used there)
        const int pa[3] = {0, 1, 2};
        const int pb[3] = {1, 2, 0};
#pragma unroll
        for (int o = 0; o < 3; o++) {
            cpx<T> p[8];
#pragma unroll
            for (int s = 0; s < 8; s++) p[s] = mk<T>(u[pa[o]][s].x * u[pb[o]][s].x, u[pa[o]][s].y * u[pb[o]][s].y);
            emit(p, 2 + o);
        }
#pragma unroll
        for (int o = 0; o < 2; o++) {
            cpx<T> p[8];
#pragma unroll
            for (int s = 0; s < 8; s++)
                p[s] = mk<T>(u[o][s].x * u[o][s].x - u[2][s].x * u[2][s].x,
                             u[o][s].y * u[o][s].y - u[2][s].y * u[2][s].y);
            emit(p, o);
        }
    } else if constexpr (!GIVER) {
        const int pa[6] = {0, 1, 2, 0, 1, 2};
        const int pb[6] = {0, 1, 2, 1, 2, 0};
#pragma unroll
        for (int o = 0; o < 6; o++) {
            cpx<T> p[8];
#pragma unroll
            for (int s = 0; s < 8; s++) p[s] = mk<T>(u[pa[o]][s].x * u[pb[o]][s].x, u[pa[o]][s].y * u[pb[o]][s].y);
            emit(p, o);
        }
    } else if constexpr (FIVE) {
        // w_z first: P_zz = u_z w_z is kept, in this thread's own LDS slots
        // (registers would cost the FP64 instance half its occupancy), for the
        // two shifted diagonal products
        __shared__ cpx<T> dl[LPB * M];
        cpx<T> *D = dl + li * M;
#pragma unroll 1
        for (int ii = 0; ii < 3; ii++) {
            const int i = ii == 0 ? 2 : ii - 1;
            cpx<T> w[8];
            zline_c2r<T, M>(w, (const cpx<T> *)a.in[3 + i] + lsafe * a.g.nzp, L, twm, twn, j);
            line_sync<M / 8>();
            if (ii == 0) {
#pragma unroll
                for (int s = 0; s < 8; s++) D[j + TPL * s] = mk<T>(u[2][s].x * w[s].x, u[2][s].y * w[s].y);
            }
#pragma unroll
            for (int jj = 0; jj < 3; jj++) {
                if (jj == i && i == 2) continue;
                cpx<T> p[8];
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    p[s] = mk<T>(u[jj][s].x * w[s].x, u[jj][s].y * w[s].y);
                    if (jj == i) p[s] = p[s] - D[j + TPL * s];
                }
                emit(p, 3 * i + jj);
            }
        }
    } else {
#pragma unroll 1
        for (int i = 0; i < 3; i++) {
            cpx<T> w[8];
            zline_c2r<T, M>(w, (const cpx<T> *)a.in[3 + i] + lsafe * a.g.nzp, L, twm, twn, j);
            line_sync<M / 8>();
#pragma unroll
            for (int jj = 0; jj < 3; jj++) {
                cpx<T> p[8];
#pragma unroll
                for (int s = 0; s < 8; s++) p[s] = mk<T>(u[jj][s].x * w[s].x, u[jj][s].y * w[s].y);
                emit(p, 3 * i + jj);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused x pass of the nonlinear term and the Lawson RK4 stage.
// ---------------------------------------------------------------------------
struct XnlArgs {
    void *P[9];
    double *V[3];
    double *acc[3];  // stage accumulator, FP64 (padded work layout)
    void *U[3];
    double *out[3];
    const double *slots_in;
    double *slots_out;
    const void *tw;
    Grid g;
    Band b;
    RkParams rk;
};

__device__ double band_alpha(const Grid &g, const Band &b, const double *slots)
{
    // alpha = eps / (2 E_band) with E_band = sum w |V|^2 over the band modes,
    // w = 1/2 on kz = 0 (Get_local_energy weights, FFF_pencil_energy.cc:59-67)
    const int n = band_nslots(b.kb);
    double e = 0.0;
    for (int s = 0; s < n; s++) {
        const int kz = s % (b.kb + 1);
        const int ky = (s / (b.kb + 1)) % (2 * b.kb + 1) - b.kb;
        const int kx = s / ((b.kb + 1) * (2 * b.kb + 1)) - b.kb;
        const double K = kmag(g, kx, ky, kz);
        if (K > b.r_in && K <= b.r_out) {
            const double w = kz == 0 ? 0.5 : 1.0;
            for (int c = 0; c < 3; c++) {
                const double re = slots[2 * (s * 3 + c)], im = slots[2 * (s * 3 + c) + 1];
                e += w * (re * re + im * im);
            }
        }
    }
    return e > 0.0 ? b.eps / (2.0 * e) : 0.0;
}

// Random band forcing of one mode (kind 2; compute_force_random.cc:65-130,
// force_amp_phase.cc:119-170, FFF_pencil_inline.h:234-248): helical
// amplitudes sqrt((e_k +- h_k/K) / tau) with phases 2 pi U(seed, 2m),
// 2 pi U(seed, 2m + 1), m = (lx Ny + ly)(Nz/2+1) + lz; U1 = i (F- - F+),
// U2 = F- + F+ in the Craya basis; Cartesian components about z.
__device__ void random_force_mode(const Grid &g, const Band &b, int lx, int ly, int lz, int kx, int ky, double K,
                                  cpx<double> *F)
{
    const int i = (int)ceil(K);
    const double e = b.shell_eh[2 * i], h = b.shell_eh[2 * i + 1];
    const double ap = sqrt((e + h / K) / b.tau), am = sqrt((e - h / K) / b.tau);
    const uint64_t m = ((uint64_t)lx * g.ny + ly) * g.nzc + lz;
    const double twopi = 6.28318530717958647692;
    const double php = twopi * uniform01(b.seed, 2 * m), phm = twopi * uniform01(b.seed, 2 * m + 1);
    const cpx<double> Fp = mk<double>(ap * cos(php), ap * sin(php)), Fm = mk<double>(am * cos(phm), am * sin(phm));
    const cpx<double> d = Fm - Fp;
    const cpx<double> U1 = mk<double>(-d.y, d.x), U2 = Fm + Fp;
    const double th = aniso_polar(g, kx, ky, lz, 3), ph = aniso_azimuth(g, kx, ky, lz, 3);
    const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    F[0] = sp * U1 + (ct * cp) * U2;
    F[1] = (-cp) * U1 + (sp * ct) * U2;
    F[2] = (-st) * U2;
}

// Order in which the product spectra are folded into N, following the
// reference's diag/offdiag accumulation (compute_diag_offdiag.cc:47-112 for
// NS, :329-414 for the giver form with P[3i+j] = FT[u_j w_i]).
static __constant__ int kNsOrder[6] = {0, 1, 2, 3, 4, 5};
// NS products xx, yy, zz, xy, yz, zx: derivative applied for N_0, N_1, N_2
// (Compute_nlin_diag, then Compute_nlin_offdiag's Add_*deriv calls).
static __constant__ int kNsSel[6][3] = {{1, 0, 0}, {0, 2, 0}, {0, 0, 3}, {2, 1, 0}, {0, 3, 2}, {3, 0, 1}};
// The trace-shifted five (xx - zz, yy - zz, xy, yz, zx; projected stages only).
static __constant__ int kNsSel5[5][3] = {{1, 0, 0}, {0, 2, 0}, {2, 1, 0}, {0, 3, 2}, {3, 0, 1}};
static __constant__ int kGiverOrder[9] = {0, 4, 8, 3, 1, 7, 5, 2, 6};

template <typename T, int L, int CB, int MODE>
__global__ void __launch_bounds__(CB *(L / 8), 2) k_x_nlin(XnlArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    // two FP64 tile buffers (separate LDS objects, so that the compiler can
    // tell accesses to one from the direct-to-LDS loads into the other): A
    // hosts the FFT exchanges and stages state inputs; B receives the next
    // product while the current one is transformed, then holds the
    // state/accumulator tile of the RK update
    __shared__ __attribute__((aligned(16))) char rawA[CB * LP * sizeof(cpx<double>)];
    __shared__ __attribute__((aligned(16))) char rawB[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    // twiddles are read from LDS: kept out of registers across the product loop
    // (1024-point lines: from global memory, L1-resident, so that two
    // blocks fit the LDS of a CU)
    constexpr bool TWL = L < 1024;
    __shared__ cpx<T> twl[TWL ? L : 1];
    cpx<double> *ldd = (cpx<double> *)rawA;
    cpx<T> *ldt = (cpx<T> *)rawA;
    cpx<double> *ldb = (cpx<double> *)rawB;
    const Grid &g = a.g;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= x_tiles(g, CB, false)) return;
    tile_columns<CB>(colw, cols, g, false, tile);
    lds_barrier();
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int cw = colw[c];
    const int ly = cw >= 0 ? cw / g.nzp : 0, lz = cw >= 0 ? cw % g.nzp : 0;
    const int ky = kwave(ly, g.ny);
    const double Ky = ky * g.kf[1], Kz = lz * g.kf[2];
    if constexpr (TWL)
        for (int i = threadIdx.x; i < L; i += CB * TPL) twl[i] = ((const cpx<T> *)a.tw)[i];
    const cpx<T> *tw = TWL ? (const cpx<T> *)twl : (const cpx<T> *)a.tw;
    const int ncw = g.ny * g.nzp, ncs = g.ny * g.nzc;
    const T sc = (T)a.rk.scale;
    const T Kyt = (T)Ky, Kzt = (T)Kz;

    // --- forward x-FFT of the products and the divergence-form sum ---
    // Product p adds i K_d FT[p] to N_i for the (i, d) pairs of the
    // divergence form; sel_i = 0 (none), 1 (Kx), 2 (Ky), 3 (Kz).  Summing
    // from zero in this order reproduces the reference's accumulation order.
    constexpr bool GIVER = MODE != 0;
    cpx<T> N[3][8];
#pragma unroll
    for (int s = 0; s < 8; s++) N[0][s] = N[1][s] = N[2][s] = mk<T>(0, 0);
    // five trace-shifted products when the nonlinear term is only used
    // projected (its gradient part i K F[u_z^2] drops out)
    const bool tf = !GIVER && a.rk.trace_free;
    const int nprod = GIVER ? 9 : (tf ? 5 : 6);
    constexpr bool DMA = tile_dma_ok<T, L, CB>();
    const int cw0 = tile * CB;  // first padded work column of the tile
    if constexpr (DMA) dma_tile<T, L, CB>((const cpx<T> *)a.P[GIVER ? kGiverOrder[0] : kNsOrder[0]], cw0, ncw, rawB);
#pragma unroll 1
    for (int pi = 0; pi < nprod; pi++) {
        const int p = GIVER ? kGiverOrder[pi] : kNsOrder[pi];
        int sel0, sel1, sel2;
        if (MODE == 1) {
            // P[3i+j] = FT[u_j w_i] contributes i K_j to N_i
            sel0 = p / 3 == 0 ? p % 3 + 1 : 0;
            sel1 = p / 3 == 1 ? p % 3 + 1 : 0;
            sel2 = p / 3 == 2 ? p % 3 + 1 : 0;
        } else if (MODE == 2) {
            // transposed: P[3i+j] contributes i K_i to N_j (MHD: N-_j = sum_i d_i(Z+_i Z-_j))
            sel0 = p % 3 == 0 ? p / 3 + 1 : 0;
            sel1 = p % 3 == 1 ? p / 3 + 1 : 0;
            sel2 = p % 3 == 2 ? p / 3 + 1 : 0;
        } else if (tf) {
            sel0 = kNsSel5[p][0];
            sel1 = kNsSel5[p][1];
            sel2 = kNsSel5[p][2];
        } else {
            sel0 = kNsSel[p][0];
            sel1 = kNsSel[p][1];
            sel2 = kNsSel[p][2];
        }
        // the line index, hidden from loop-invariant code motion: the LDS
        // addresses of the FFT passes are recomputed each product (a few
        // integer ops) rather than held in registers across the loop, where
        // their spill reloads would wait for the in-flight prefetch
        int jj = j;
        asm volatile("" : "+v"(jj));
        cpx<T> v[8];
        if constexpr (DMA) {
            // product p has landed in B; read it, then start the next one
            dma_wait();
            lds_barrier();
#pragma unroll
            for (int s = 0; s < 8; s++) v[s] = dma_tile_read<T, L, CB>(rawB, jj + TPL * s, c);
            lds_barrier();
            if (pi + 1 < nprod)
                dma_tile<T, L, CB>((const cpx<T> *)a.P[GIVER ? kGiverOrder[pi + 1] : kNsOrder[pi + 1]], cw0, ncw,
                                   rawB);
            else if (a.rk.stage > 0)
                // the state of the first component, for the RK update below
                dma_state_tile<L, CB>((const cpx<double> *)a.V[0], cols, ncs, rawB);
        } else {
            stage_in<L, CB>(ldt, (const cpx<T> *)a.P[p], colw, ncw);
            lds_barrier();
#pragma unroll
            for (int s = 0; s < 8; s++) v[s] = ldt[c * LP + lpad(j + TPL * s)];
            lds_barrier();
        }
        fft_line<L, -1>(v, ldt + c * LP, tw, jj);
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const T Kx = (T)(kwave(jj + TPL * s, g.nx) * g.kf[0]);
            const cpx<T> q = sc * v[s];
            const T K0 = sel0 == 1 ? Kx : (sel0 == 2 ? Kyt : (sel0 == 3 ? Kzt : (T)0));
            const T K1 = sel1 == 1 ? Kx : (sel1 == 2 ? Kyt : (sel1 == 3 ? Kzt : (T)0));
            const T K2 = sel2 == 1 ? Kx : (sel2 == 2 ? Kyt : (sel2 == 3 ? Kzt : (T)0));
            N[0][s] = N[0][s] + mul_i(K0, q);
            N[1][s] = N[1][s] + mul_i(K1, q);
            N[2][s] = N[2][s] + mul_i(K2, q);
        }
        if constexpr (!DMA) lds_barrier();
    }
    // the FFT exchanges in A are done before A and B are reused below (B may
    // be receiving the first state component)
    if constexpr (DMA) lds_barrier_async();

    const int stage = a.rk.stage;
    const bool want_rhs = stage > 0 || a.rk.write_rhs;
    // --- force, pressure projection, negation (Compute_force_TO_rhs), in
    // place on N: N <- -(N - F + i K p), p = i K.(N - F) / K^2 (computed in
    // FP64, stored in the transform precision) ---
    double alpha = 0.0;
    bool alpha_done = false;
    if (want_rhs) {
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int lx = j + TPL * s;
            const int kx = kwave(lx, g.nx);
            const double Kx = kx * g.kf[0];
            cpx<double> n0 = cvt<double>(N[0][s]), n1 = cvt<double>(N[1][s]), n2 = cvt<double>(N[2][s]);
            if (a.b.kind == 1 && cw >= 0) {
                const int sl = band_slot(a.b, kx, ky, lz);
                if (sl >= 0) {
                    const double K = kmag(g, kx, ky, lz);
                    if (K > a.b.r_in && K <= a.b.r_out) {
                        if (!alpha_done) { alpha = band_alpha(g, a.b, a.slots_in); alpha_done = true; }
                        const cpx<double> *sv = (const cpx<double> *)a.slots_in + sl * 3;
                        n0 = n0 - alpha * sv[0];
                        n1 = n1 - alpha * sv[1];
                        n2 = n2 - alpha * sv[2];
                    }
                }
            } else if (a.b.kind == 2 && cw >= 0) {
                const double K = kmag(g, kx, ky, lz);
                if (K > a.b.r_in && K <= a.b.r_out) {
                    cpx<double> F[3];
                    random_force_mode(g, a.b, lx, ly, lz, kx, ky, K, F);
                    n0 = n0 - F[0];
                    n1 = n1 - F[1];
                    n2 = n2 - F[2];
                }
            }
            cpx<double> div = mul_i(Kx, n0);
            div = div + mul_i(Ky, n1);
            div = div + mul_i(Kz, n2);
            const double kxs = Kx * Kx;
            const double ks = (kxs + Ky * Ky) + Kz * Kz;
            cpx<double> pr = mk<double>(0, 0);
            if (!(lx == 0 && ly == 0 && lz == 0)) pr = mk<double>(div.x / ks, div.y / ks);
            n0 = n0 + mul_i(Kx, pr);
            n1 = n1 + mul_i(Ky, pr);
            n2 = n2 + mul_i(Kz, pr);
            N[0][s] = cvt<T>(mk<double>(-n0.x, -n0.y));
            N[1][s] = cvt<T>(mk<double>(-n1.x, -n1.y));
            N[2][s] = cvt<T>(mk<double>(-n2.x, -n2.y));
        }
    }
#define TG_TAKE(comp, nc)                                                           \
    do {                                                                            \
        if ((comp) == 0) {                                                          \
            _Pragma("unroll") for (int s_ = 0; s_ < 8; s_++) nc[s_] = N[0][s_];    \
        } else if ((comp) == 1) {                                                   \
            _Pragma("unroll") for (int s_ = 0; s_ < 8; s_++) nc[s_] = N[1][s_];    \
        } else {                                                                    \
            _Pragma("unroll") for (int s_ = 0; s_ < 8; s_++) nc[s_] = N[2][s_];    \
        }                                                                           \
    } while (0)

    if (stage == 0) {
        // Compute_nlin (and optionally Compute_rhs) output, FP64 boundary layout
#pragma unroll 1
        for (int comp = 0; comp < 3; comp++) {
            cpx<T> nc[8];
            TG_TAKE(comp, nc);
#pragma unroll
            for (int s = 0; s < 8; s++) ldd[c * LP + lpad(j + TPL * s)] = cvt<double>(nc[s]);
            lds_barrier();
            stage_out<L, CB>(ldd, (cpx<double> *)a.out[comp], cols, ncs);
            lds_barrier();
        }
        return;
    }

    // --- Lawson RK4 stage algebra (RK.cc:91-170, Tadvance_functions.cc:298-324,566-570) ---
    // V: FP64 state, boundary layout; acc: FP64, padded work layout (the RK
    // state accumulation stays FP64 whatever the transform precision).
    const double dt = a.rk.dt;
    const double c_half = 0.5 * dt, c_one = 1.0 * dt;
    const double c_sixth = (1.0 / 6) * dt, c_third = (1.0 / 3) * dt;
    if constexpr (DMA) {
        // Component K: its state V[K] sits in tile X (direct-to-LDS load,
        // column-major chunks), the accumulator is staged into tile Y, which
        // then takes the updated accumulator (or state); the next stage input
        // is x-inverted with X as the exchange buffer, and V[K+1] is loaded
        // into Y while that input is stored.  The components are unrolled so
        // that X and Y are known LDS objects (the waits for the direct-to-LDS
        // loads then cover only the tile they write).
        auto rk_comp = [&](auto kc, char *X, char *Y) {
            constexpr int K = decltype(kc)::value;
            cpx<double> *Yd = (cpx<double> *)Y;
            cpx<T> *Xt = (cpx<T> *)X;
            if (stage >= 2) stage_in<L, CB>(Yd, (const cpx<double> *)a.acc[K], colw, ncw);
            dma_wait();
            lds_barrier();
            cpx<T> v[8];
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int lx = j + TPL * s;
                const int li = c * LP + lpad(lx);
                const double Kx = kwave(lx, g.nx) * g.kf[0];
                const double kxs = Kx * Kx;
                const double ks = (kxs + Ky * Ky) + Kz * Kz;
                const double Eh = visc_exp(a.rk.fh, a.rk.fhh, a.rk.hq, ks);
                const cpx<double> r = cvt<double>(N[K][s]);
                const cpx<double> v0 = dma_tile_read<double, L, CB>(X, lx, c);
                cpx<double> vn, ac;
                if (stage == 1) {
                    cpx<double> d0 = v0;
                    if (dealias_zero(g, lx, ly, lz)) d0 = mk<double>(0, 0);
                    const double E1 = visc_exp(a.rk.f1, a.rk.f1h, a.rk.hq, ks);
                    ac = c_sixth * (E1 * r);
                    vn = Eh * (d0 + c_half * r);
                } else if (stage == 2) {
                    ac = Yd[li] + c_third * (Eh * r);
                    vn = Eh * v0 + c_half * r;
                } else if (stage == 3) {
                    ac = Yd[li] + c_third * (Eh * r);
                    vn = Eh * (Eh * v0 + c_one * r);
                } else {
                    ac = Yd[li] + c_sixth * r;
                    const double E1 = visc_exp(a.rk.f1, a.rk.f1h, a.rk.hq, ks);
                    vn = E1 * v0 + ac;
                }
                Yd[li] = mk<double>(stage == 4 ? vn.x : ac.x, stage == 4 ? vn.y : ac.y);
                if (stage == 4 && !(isfinite(vn.x) && isfinite(vn.y)) && a.rk.nonfinite) *a.rk.nonfinite = 1;
                if (a.b.kind == 1 && cw >= 0) {
                    const int kx = kwave(lx, g.nx);
                    const int sl = band_slot(a.b, kx, ky, lz);
                    if (sl >= 0) {
                        const double Kb = kmag(g, kx, ky, lz);
                        if (Kb > a.b.r_in && Kb <= a.b.r_out) ((cpx<double> *)a.slots_out)[sl * 3 + K] = vn;
                    }
                }
                if (dealias_zero(g, lx, ly, lz)) vn = mk<double>(0, 0);
                v[s] = cvt<T>(vn);
            }
            lds_barrier();
            if (stage == 4) stage_out<L, CB>(Yd, (cpx<double> *)a.V[K], cols, ncs);
            else stage_out<L, CB>(Yd, (cpx<double> *)a.acc[K], colw, ncw);
            fft_line<L, +1>(v, Xt + c * LP, tw, j);
#pragma unroll
            for (int s = 0; s < 8; s++) Xt[c * LP + lpad(j + TPL * s)] = v[s];
            lds_barrier();
            if constexpr (K < 2) dma_state_tile<L, CB>((const cpx<double> *)a.V[K + 1], cols, ncs, Y);
            stage_out<L, CB>(Xt, (cpx<T> *)a.U[K], colw, ncw);
            lds_barrier_async();
        };
        rk_comp(std::integral_constant<int, 0>{}, rawB, rawA);
        rk_comp(std::integral_constant<int, 1>{}, rawA, rawB);
        rk_comp(std::integral_constant<int, 2>{}, rawB, rawA);
        return;
    }
#pragma unroll 1
    for (int comp = 0; comp < 3; comp++) {
        cpx<T> nc[8];
        TG_TAKE(comp, nc);
        stage_in<L, CB>(ldd, (const cpx<double> *)a.V[comp], cols, ncs);
        if (stage >= 2) stage_in<L, CB>(ldb, (const cpx<double> *)a.acc[comp], colw, ncw);
        lds_barrier();
        cpx<T> v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int lx = j + TPL * s;
            const int li = c * LP + lpad(lx);
            const double Kx = kwave(lx, g.nx) * g.kf[0];
            const double kxs = Kx * Kx;
            const double ks = (kxs + Ky * Ky) + Kz * Kz;
            const double Eh = visc_exp(a.rk.fh, a.rk.fhh, a.rk.hq, ks);
            const cpx<double> r = cvt<double>(nc[s]);
            const cpx<double> v0 = ldd[li];
            cpx<double> vn, ac;
            if (stage == 1) {
                cpx<double> d0 = v0;
                if (dealias_zero(g, lx, ly, lz)) d0 = mk<double>(0, 0);
                const double E1 = visc_exp(a.rk.f1, a.rk.f1h, a.rk.hq, ks);
                ac = c_sixth * (E1 * r);
                vn = Eh * (d0 + c_half * r);
            } else if (stage == 2) {
                ac = ldb[li] + c_third * (Eh * r);
                vn = Eh * v0 + c_half * r;
            } else if (stage == 3) {
                ac = ldb[li] + c_third * (Eh * r);
                vn = Eh * (Eh * v0 + c_one * r);
            } else {
                ac = ldb[li] + c_sixth * r;
                const double E1 = visc_exp(a.rk.f1, a.rk.f1h, a.rk.hq, ks);
                vn = E1 * v0 + ac;
            }
            // accumulator (stages 1-3) or the new state (stage 4) back to B
            // (component-wise select: selecting the objects keeps them in memory)
            ldb[li] = mk<double>(stage == 4 ? vn.x : ac.x, stage == 4 ? vn.y : ac.y);
            if (stage == 4 && !(isfinite(vn.x) && isfinite(vn.y)) && a.rk.nonfinite) *a.rk.nonfinite = 1;
            // next stage input: band slots, dealias
            if (a.b.kind == 1 && cw >= 0) {
                const int kx = kwave(lx, g.nx);
                const int sl = band_slot(a.b, kx, ky, lz);
                if (sl >= 0) {
                    const double K = kmag(g, kx, ky, lz);
                    if (K > a.b.r_in && K <= a.b.r_out) ((cpx<double> *)a.slots_out)[sl * 3 + comp] = vn;
                }
            }
            if (dealias_zero(g, lx, ly, lz)) vn = mk<double>(0, 0);
            v[s] = cvt<T>(vn);
        }
        lds_barrier();
        if (stage == 4) stage_out<L, CB>(ldb, (cpx<double> *)a.V[comp], cols, ncs);
        else stage_out<L, CB>(ldb, (cpx<double> *)a.acc[comp], colw, ncw);
        // inverse x-FFT of the next stage input (exchanges in buffer A)
        fft_line<L, +1>(v, ldt + c * LP, tw, j);
#pragma unroll
        for (int s = 0; s < 8; s++) ldt[c * LP + lpad(j + TPL * s)] = v[s];
        lds_barrier();
        stage_out<L, CB>(ldt, (cpx<T> *)a.U[comp], colw, ncw);
        lds_barrier();
    }
}
#undef TG_TAKE

#if TG_LEN_UNIT == 0
// ---------------------------------------------------------------------------
// Strong reality condition (FFF_pencil_basic.cc:139-179) and dealias.
// ---------------------------------------------------------------------------
struct PlaneArgs {
    double *V[3];
    Grid g;
    int ncomp;
};

__global__ void k_reality(PlaneArgs a)
{
    const Grid &g = a.g;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nx * g.ny) return;
    const int lx = idx / g.ny, ly = idx % g.ny;
    const int ky = kwave(ly, g.ny);
    for (int comp = 0; comp < a.ncomp; comp++) {
        cpx<double> *A = (cpx<double> *)a.V[comp];
        const size_t t = ((size_t)lx * g.ny + ly) * g.nzc;
        int src = -1;
        if (lx > g.nx / 2 && ky != g.ny / 2) {
            const int mkx = -kwave(lx, g.nx);
            const int mky = -ky;
            const int iy = mky >= 0 ? mky : mky + g.ny;
            src = mkx * g.ny + iy;
        } else if (lx == 0 && ky < 0) {
            src = -ky;
        }
        if (src >= 0) A[t] = conj(A[(size_t)src * g.nzc]);
        A[t + g.nzc - 1] = mk<double>(0, 0);
    }
}

__global__ void k_dealias(PlaneArgs a)
{
    const Grid &g = a.g;
    const size_t n = (size_t)g.nx * g.ny * g.nzc;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int lz = i % g.nzc;
        const int ly = (i / g.nzc) % g.ny;
        const int lx = i / ((size_t)g.nzc * g.ny);
        if (dealias_zero(g, lx, ly, lz))
            for (int comp = 0; comp < a.ncomp; comp++) ((cpx<double> *)a.V[comp])[i] = mk<double>(0, 0);
    }
}

#endif  // TG_LEN_UNIT == 0

// ---------------------------------------------------------------------------
// Per-length launch functions (instantiated in the length units) and the
// launchers with their size dispatch.
// ---------------------------------------------------------------------------
template <typename T, int LL>
hipError_t colfft_len(const ColArgs &a, int dir, int nfields, hipStream_t st)
{
    // at most 512 threads a block, except FP32 lines of 512 points and
    // more: 1024-thread blocks (16 columns at 512 points, 8 at 1024) run
    // 2% and 9% faster than the 512-thread tiles of half the width
    constexpr int CB = std::min(col_block<T>(LL, sizeof(cpx<T>)),
                                std::max(1, (sizeof(T) == 4 && LL >= 512 ? 1024 : 512) / (LL / 8)));
    dim3 grid(x_grid(a.nouter * ((a.ncol + CB - 1) / CB)), nfields);
    if (dir < 0) hipLaunchKernelGGL((k_colfft<T, LL, -1, CB>), grid, dim3(CB * (LL / 8)), 0, st, a);
    else hipLaunchKernelGGL((k_colfft<T, LL, +1, CB>), grid, dim3(CB * (LL / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int LL>
hipError_t xinv_state_len(const XinvArgs &a, hipStream_t st)
{
    constexpr int CB = col_block<T>(LL, sizeof(cpx<double>));
    dim3 grid(x_grid(x_tiles(a.g, CB, (a.flags & 4) != 0)));
    hipLaunchKernelGGL((k_xinv_state<T, LL, CB>), grid, dim3(CB * (LL / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int LL>
hipError_t xfwd_spec_len(const XfwdArgs &a, hipStream_t st)
{
    constexpr int CB = col_block<T>(LL, sizeof(cpx<double>));
    dim3 grid(x_grid(x_tiles(a.g, CB, false)));
    hipLaunchKernelGGL((k_xfwd_spec<T, LL, CB>), grid, dim3(CB * (LL / 8)), 0, st, a);
    return hipGetLastError();
}

template <int MM>
inline dim3 z_grid(const Grid &g)
{
    constexpr int LPB = z_lines_per_block<MM>();
    const size_t nlines = (size_t)g.nx * g.ny;
    return dim3((unsigned)((nlines + LPB - 1) / LPB));
}

template <typename T, int MM>
hipError_t z_c2r_len(const ZArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL((k_z_c2r<T, MM>), z_grid<MM>(a.g), dim3(z_lines_per_block<MM>() * (MM / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int MM>
hipError_t z_r2c_len(const ZArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL((k_z_r2c<T, MM>), z_grid<MM>(a.g), dim3(z_lines_per_block<MM>() * (MM / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int MM>
hipError_t z_products_len(const ZArgs &a, int giver, hipStream_t st)
{
    const dim3 grid = z_grid<MM>(a.g), block(z_lines_per_block<MM>() * (MM / 8));
    if (giver && a.ncomp == 8) hipLaunchKernelGGL((k_z_products<T, MM, true, true>), grid, block, 0, st, a);
    else if (giver) hipLaunchKernelGGL((k_z_products<T, MM, true>), grid, block, 0, st, a);
    else if (a.ncomp == 5) hipLaunchKernelGGL((k_z_products<T, MM, false, true>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_z_products<T, MM, false>), grid, block, 0, st, a);
    return hipGetLastError();
}

#define TG_XNL(CBV, M) \
    hipLaunchKernelGGL((k_x_nlin<T, LL, CBV, M>), dim3(x_grid(x_tiles(a.g, CBV, false))), dim3(CBV * (LL / 8)), 0, st, a)

template <typename T, int LL>
hipError_t x_nlin_len(const XnlArgs &a, int mode, int x_tile_columns, hipStream_t st)
{
    // two FP64 tile buffers per block: the LDS-budget width (two blocks
    // per CU), or twice that width (one block per CU)
    constexpr int CB = col_block<T>(LL, 2 * sizeof(cpx<double>));
    constexpr int CBW = 2 * CB * (LL / 8) <= 1024 ? 2 * CB : CB;
    const bool wide = x_tile_columns == CBW && CBW != CB;
    if (wide) {
        if (mode == 1) TG_XNL(CBW, 1);
        else if (mode == 2) TG_XNL(CBW, 2);
        else TG_XNL(CBW, 0);
    } else {
        if (mode == 1) TG_XNL(CB, 1);
        else if (mode == 2) TG_XNL(CB, 2);
        else TG_XNL(CB, 0);
    }
    return hipGetLastError();
}
#undef TG_XNL

#define TG_KERNELS_L(N, T, P)                                                                  \
    P template hipError_t colfft_len<T, N>(const ColArgs &, int, int, hipStream_t);           \
    P template hipError_t xinv_state_len<T, N>(const XinvArgs &, hipStream_t);                \
    P template hipError_t xfwd_spec_len<T, N>(const XfwdArgs &, hipStream_t);                 \
    P template hipError_t x_nlin_len<T, N>(const XnlArgs &, int, int, hipStream_t);
#define TG_KERNELS_M(N, T, P)                                                                  \
    P template hipError_t z_c2r_len<T, N>(const ZArgs &, hipStream_t);                        \
    P template hipError_t z_r2c_len<T, N>(const ZArgs &, hipStream_t);                        \
    P template hipError_t z_products_len<T, N>(const ZArgs &, int, hipStream_t);

#if TG_LEN_UNIT == 0
TG_ALL_UNITS(TG_KERNELS_L, TG_KERNELS_M, extern)

template <typename T>
hipError_t launch_colfft(const Grid &g, int axis, int dir, void *const *fields, int nfields, const void *tw,
                         hipStream_t st)
{
    ColArgs a{};
    for (int i = 0; i < nfields; i++) a.f[i] = fields[i];
    a.tw = tw;
    int L;
    if (axis == 1) {
        L = g.ny;
        a.nouter = g.nx;
        a.outer_stride = (long long)g.ny * g.nzp;
        a.row_stride = g.nzp;
        a.ncol = g.nzc;
    } else {
        L = g.nx;
        a.nouter = 1;
        a.outer_stride = 0;
        a.row_stride = (long long)g.ny * g.nzp;
        a.ncol = g.ny * g.nzp;
    }
    TG_DISPATCH_L(T, L, return (colfft_len<T, LL>)(a, dir, nfields, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_xinv_state(const Grid &g, const Band &b, double *const *V, int ncomp, void *const *U,
                             double *slots, int flags, const void *twx, hipStream_t st, double r_lo, double r_hi,
                             double a_lo, double a_hi, int dirn)
{
    XinvArgs a{};
    a.r_lo = r_lo;
    a.r_hi = r_hi;
    a.a_lo = a_lo;
    a.a_hi = a_hi;
    a.dirn = dirn;
    for (int i = 0; i < ncomp; i++) { a.V[i] = V[i]; a.U[i] = U[i]; }
    a.slots = slots;
    a.tw = twx;
    a.g = g;
    a.b = b;
    a.ncomp = ncomp;
    a.flags = flags;
    TG_DISPATCH_L(T, g.nx, return (xinv_state_len<T, LL>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_xfwd_spec(const Grid &g, void *const *W, int ncomp, double *const *out, double scale,
                            const void *twx, hipStream_t st)
{
    XfwdArgs a{};
    for (int i = 0; i < ncomp; i++) { a.W[i] = W[i]; a.out[i] = out[i]; }
    a.tw = twx;
    a.g = g;
    a.scale = scale;
    a.ncomp = ncomp;
    TG_DISPATCH_L(T, g.nx, return (xfwd_spec_len<T, LL>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_z_c2r(const Grid &g, void *const *W, int ncomp, double *const *R, const void *twm, const void *twn,
                        hipStream_t st)
{
    ZArgs a{};
    for (int i = 0; i < ncomp; i++) { a.in[i] = W[i]; a.out[i] = R[i]; }
    a.twm = twm;
    a.twn = twn;
    a.g = g;
    a.ncomp = ncomp;
    TG_DISPATCH_M(T, g.nz / 2, return (z_c2r_len<T, MM>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_z_r2c(const Grid &g, const double *const *R, int ncomp, void *const *W, const void *twm,
                        const void *twn, hipStream_t st)
{
    ZArgs a{};
    for (int i = 0; i < ncomp; i++) { a.in[i] = (void *)R[i]; a.out[i] = W[i]; }
    a.twm = twm;
    a.twn = twn;
    a.g = g;
    a.ncomp = ncomp;
    TG_DISPATCH_M(T, g.nz / 2, return (z_r2c_len<T, MM>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_z_products(const Grid &g, void *const *U, void *const *P, int giver, const void *twm,
                             const void *twn, hipStream_t st, void *const *tape)
{
    ZArgs a{};
    // giver 2: the NS form with the five trace-shifted products; 3: the giver
    // form with the eight trace-shifted ones (MHD stage)
    a.ncomp = giver == 2 ? 5 : (giver == 3 ? 8 : (giver == 1 ? 9 : 6));
    giver = giver == 1 || giver == 3;
    if (tape && !giver)
        for (int i = 0; i < 3; i++) a.tape[i] = tape[i];
    for (int i = 0; i < (giver ? 6 : 3); i++) a.in[i] = U[i];
    for (int i = 0; i < a.ncomp; i++) a.out[i] = P[i];
    a.twm = twm;
    a.twn = twn;
    a.g = g;
    TG_DISPATCH_M(T, g.nz / 2, return (z_products_len<T, MM>)(a, giver, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_x_nlin(const Grid &g, const Band &b, const RkParams &rk, void *const *P, double *const *V,
                         double *const *acc, void *const *U, const double *slots_in, double *slots_out,
                         double *const *out, const void *twx, hipStream_t st, int x_tile_columns)
{
    XnlArgs a{};
    // rk.stage -1: giver form, -2: transposed giver form (nonlinear term only)
    const int giver = rk.stage == -1 ? 1 : (rk.stage == -2 ? 2 : 0);
    for (int i = 0; i < (giver ? 9 : 6); i++) a.P[i] = P[i];
    for (int i = 0; i < 3; i++) {
        a.V[i] = V ? V[i] : nullptr;
        a.acc[i] = acc ? acc[i] : nullptr;
        a.U[i] = U ? U[i] : nullptr;
        a.out[i] = out ? out[i] : nullptr;
    }
    a.slots_in = slots_in;
    a.slots_out = slots_out;
    a.tw = twx;
    a.g = g;
    a.b = b;
    a.rk = rk;
    if (giver) {
        a.rk.stage = 0;
        a.rk.write_rhs = 0;
    }
    TG_DISPATCH_L(T, g.nx, return (x_nlin_len<T, LL>)(a, giver, x_tile_columns, st));
    return hipErrorInvalidValue;
}

__global__ void k_fill_int(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_int(int *p, int n, int v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_int, dim3((n + 63) / 64), dim3(64), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_reality(const Grid &g, double *const *V, int ncomp, hipStream_t st)
{
    PlaneArgs a{};
    for (int i = 0; i < ncomp; i++) a.V[i] = V[i];
    a.g = g;
    a.ncomp = ncomp;
    const int n = g.nx * g.ny;
    hipLaunchKernelGGL(k_reality, dim3((n + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_dealias(const Grid &g, double *const *V, int ncomp, hipStream_t st)
{
    PlaneArgs a{};
    for (int i = 0; i < ncomp; i++) a.V[i] = V[i];
    a.g = g;
    a.ncomp = ncomp;
    hipLaunchKernelGGL(k_dealias, dim3(2048), dim3(256), 0, st, a);
    return hipGetLastError();
}

#define TG_INST(T)                                                                                              \
    template hipError_t launch_colfft<T>(const Grid &, int, int, void *const *, int, const void *, hipStream_t); \
    template hipError_t launch_xinv_state<T>(const Grid &, const Band &, double *const *, int, void *const *,   \
                                             double *, int, const void *, hipStream_t, double, double, double, \
                                             double, int);                                                     \
    template hipError_t launch_xfwd_spec<T>(const Grid &, void *const *, int, double *const *, double,          \
                                            const void *, hipStream_t);                                        \
    template hipError_t launch_z_c2r<T>(const Grid &, void *const *, int, double *const *, const void *,        \
                                        const void *, hipStream_t);                                            \
    template hipError_t launch_z_r2c<T>(const Grid &, const double *const *, int, void *const *, const void *,  \
                                        const void *, hipStream_t);                                            \
    template hipError_t launch_z_products<T>(const Grid &, void *const *, void *const *, int, const void *,     \
                                             const void *, hipStream_t, void *const *);                        \
    template hipError_t launch_x_nlin<T>(const Grid &, const Band &, const RkParams &, void *const *,           \
                                         double *const *, double *const *, void *const *, const double *,      \
                                         double *, double *const *, const void *, hipStream_t, int);

TG_INST(double)
TG_INST(float)
#else
TG_THIS_UNIT(TG_KERNELS_L, TG_KERNELS_M)
#endif  // TG_LEN_UNIT

}  // namespace tg
