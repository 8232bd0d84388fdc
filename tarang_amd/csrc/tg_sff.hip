// tg_sff.hip -- SFF basis kernels (sin/cos along x) and the Rayleigh-Benard
// stage pipeline on it.
//
// x transforms.  The reference's sin/cos transforms live in the absent fftk
// library; this build's convention (cell-centred samples x_j = (j+1/2)L/N,
// f_j = sum_m w_m A_m phi_m(x_j), w_0 = 1, w_m = 2, phi = sin or cos of
// m pi x / L; forward A = DCT-II/(2N), or the DST-II/(2N) shifted so that
// index m holds sin(m pi x / L), A_0 = 0) follows the reference's SFF energy
// (SFF_pencil_energy.cc:88-104) and x-derivative (SFF_pencil_tr.cc:84-105)
// formulas.  Each line transform is one complex FFT of length N
// (Makhoul's reordering): v_n = f_{2n}, v_{N-1-n} = f_{2n+1};
//   cosine forward   2N A_k = W^k V_k + W^-k V_{N-k},   V = FFT(v)
//   sine forward     the cosine transform of (-1)^j f_j, read backwards:
//                    2N A_m = i (W^m V_m - W^-m V_{N-m}), A_0 = 0
//   cosine inverse   V_k = W^-k (A_k - i A_{N-k}), v = IFFT(V) (unscaled)
//   sine inverse     V_k = W^-k (A_{N-k} - i A_k), V_0 = 0, then (-1)^j
// with W = exp(-i pi / (2N)) (table twq) and the (k, N-k) pairs exchanged
// through the LDS line buffer.  A thread holds k = j + TPL*s (s < 8), and
// every pass forms W^k = W^j W^(TPL s) from those TPL + 8 table entries
// (QuarterTw), so all sin/cos passes round alike and the fused kernels keep
// only 8 + TPL entries in LDS.  Along y and z the FFF passes are reused.
#include <type_traits>

#include "tg_device.h"
#include "tg_pass.h"
#include "tg_sff.h"

namespace tg {

// sample index whose value sits at Makhoul position n
__device__ __forceinline__ int mk_perm(int n, int N) { return n < N / 2 ? 2 * n : 2 * N - 2 * n - 1; }

template <typename T>
__device__ __forceinline__ cpx<T> times_i(cpx<T> z) { return mk<T>(-z.y, z.x); }

// Post-twiddle W^(j + TPL*s) for the sin/cos line transforms.
template <typename T>
struct QuarterTw {
    const cpx<T> *lo;  // W^j, j < TPL
    const cpx<T> *hi;  // W^(TPL s), s < 8
    __device__ __forceinline__ cpx<T> operator()(int j, int s) const { return lo[j] * hi[s]; }
};

// The core: v[s] holds the reordered (and, for sine, sign-alternated) input
// at Makhoul position j + TPL*s; `col` is a free LDS line buffer.
template <int L, typename T, typename Q>
__device__ __forceinline__ void sc_forward_core(cpx<T> *v, cpx<T> *col, const cpx<T> *tw, Q twq, int j, bool sine)
{
    constexpr int TPL = L / 8;
    fft_line<L, -1>(v, col, tw, j);
#pragma unroll
    for (int s = 0; s < 8; s++) col[lpad(j + TPL * s)] = v[s];
    line_sync<TPL>();
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int k = j + TPL * s;
        const cpx<T> vn = col[lpad(k == 0 ? 0 : L - k)];
        const cpx<T> w = twq(j, s);
        const cpx<T> p = w * v[s];
        const cpx<T> q = conj(w) * vn;
        if (!sine) v[s] = p + q;
        else v[s] = k == 0 ? mk<T>(0, 0) : times_i(p - q);
    }
    line_sync<TPL>();
}

// Makhoul input gather of sample mk_perm(n) from `load(sample)`.
template <int L, typename T, typename Load>
__device__ __forceinline__ void sc_gather(cpx<T> *v, int j, bool sine, Load load)
{
    constexpr int TPL = L / 8;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int n = j + TPL * s;
        cpx<T> x = load(mk_perm(n, L));
        if (sine && n >= L / 2) x = mk<T>(-x.x, -x.y);
        v[s] = x;
    }
}

// Forward sin/cos transform of the line held in natural order in the LDS
// line buffer `col`; on return v[s] = 2N A_k at k = j + TPL*s.  The caller
// synchronises before (col written by other threads) and after (col reused).
template <int L, typename T, typename Q>
__device__ __forceinline__ void sc_forward(cpx<T> *v, cpx<T> *col, const cpx<T> *tw, Q twq, int j, bool sine)
{
    sc_gather<L, T>(v, j, sine, [&](int r) { return col[lpad(r)]; });
    line_sync<L / 8>();
    sc_forward_core<L>(v, col, tw, twq, j, sine);
}

// Inverse: A_k in v[s] (k = j + TPL*s); on return the samples are in `col`
// in natural order (the caller synchronises before reading them).
template <int L, typename T, typename Q>
__device__ __forceinline__ void sc_inverse(cpx<T> *v, cpx<T> *col, const cpx<T> *tw, Q twq, int j, bool sine)
{
    constexpr int TPL = L / 8;
#pragma unroll
    for (int s = 0; s < 8; s++) col[lpad(j + TPL * s)] = v[s];
    line_sync<TPL>();
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int k = j + TPL * s;
        const cpx<T> an = k == 0 ? mk<T>(0, 0) : col[lpad(L - k)];
        const cpx<T> ak = v[s];
        // component-wise selects (a select of the two values as objects keeps
        // v[] out of registers)
        const cpx<T> p = mk<T>(sine ? an.x : ak.x, sine ? an.y : ak.y);
        const cpx<T> q = mk<T>(sine ? ak.x : an.x, sine ? ak.y : an.y);
        cpx<T> z = conj(twq(j, s)) * mk<T>(p.x + q.y, p.y - q.x);  // W^-k (p - i q)
        if (sine && k == 0) z = mk<T>(0, 0);
        v[s] = z;
    }
    line_sync<TPL>();
    fft_line<L, +1>(v, col, tw, j);
    line_sync<TPL>();
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int n = j + TPL * s;
        cpx<T> x = v[s];
        if (sine && n >= L / 2) x = mk<T>(-x.x, -x.y);
        col[lpad(mk_perm(n, L))] = x;
    }
}

// ---------------------------------------------------------------------------
// Plain x passes (tg_forward / tg_inverse, stage-1 inputs, reality patches)
// ---------------------------------------------------------------------------
struct XscArgs {
    double *V[4];
    void *U[4];
    const void *tw, *twq;
    Grid g;
    int ncomp, parity, flags;
    double scale;
};

template <typename T, int L, int CB>
__global__ void __launch_bounds__(CB *(L / 8)) k_xinv_sc(XscArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    __shared__ __attribute__((aligned(16))) char raw[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    cpx<double> *ldd = (cpx<double> *)raw;
    cpx<T> *ldt = (cpx<T> *)raw;
    const Grid &g = a.g;
    __shared__ cpx<T> qhi[8];
    const bool patch = (a.flags & 4) != 0;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= x_tiles(g, CB, patch)) return;
    tile_columns<CB>(colw, cols, g, patch, tile);
    if (threadIdx.x < 8) qhi[threadIdx.x] = ((const cpx<T> *)a.twq)[(L / 8) * threadIdx.x];
    lds_barrier();
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int cw = colw[c];
    const int ly = cw >= 0 ? cw / g.nzp : 0, lz = cw >= 0 ? cw % g.nzp : 0;
    const cpx<T> *tw = (const cpx<T> *)a.tw;
    const QuarterTw<T> twq{(const cpx<T> *)a.twq, qhi};
    const int ncw = g.ny * g.nzp, ncs = g.ny * g.nzc;
    for (int comp = 0; comp < a.ncomp; comp++) {
        stage_in<L, CB>(ldd, (const cpx<double> *)a.V[comp], cols, ncs);
        lds_barrier();
        cpx<T> v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int lx = j + TPL * s;
            cpx<double> d = ldd[c * LP + lpad(lx)];
            if ((a.flags & 1) && dealias_zero(g, lx, ly, lz)) d = mk<double>(0, 0);
            v[s] = cvt<T>(d);
        }
        lds_barrier();
        sc_inverse<L>(v, ldt + c * LP, tw, twq, j, ((a.parity >> comp) & 1) != 0);
        lds_barrier();
        stage_out<L, CB>(ldt, (cpx<T> *)a.U[comp], colw, ncw);
        lds_barrier();
    }
}

template <typename T, int L, int CB>
__global__ void __launch_bounds__(CB *(L / 8)) k_xfwd_sc(XscArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    __shared__ __attribute__((aligned(16))) char raw[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    cpx<double> *ldd = (cpx<double> *)raw;
    cpx<T> *ldt = (cpx<T> *)raw;
    const Grid &g = a.g;
    __shared__ cpx<T> qhi[8];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= x_tiles(g, CB, false)) return;
    tile_columns<CB>(colw, cols, g, false, tile);
    if (threadIdx.x < 8) qhi[threadIdx.x] = ((const cpx<T> *)a.twq)[(L / 8) * threadIdx.x];
    lds_barrier();
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const cpx<T> *tw = (const cpx<T> *)a.tw;
    const QuarterTw<T> twq{(const cpx<T> *)a.twq, qhi};
    const int ncw = g.ny * g.nzp, ncs = g.ny * g.nzc;
    for (int comp = 0; comp < a.ncomp; comp++) {
        stage_in<L, CB>(ldt, (const cpx<T> *)a.U[comp], colw, ncw);
        lds_barrier();
        cpx<T> v[8];
        sc_forward<L>(v, ldt + c * LP, tw, twq, j, ((a.parity >> comp) & 1) != 0);
        lds_barrier();
#pragma unroll
        for (int s = 0; s < 8; s++) ldd[c * LP + lpad(j + TPL * s)] = a.scale * cvt<double>(v[s]);
        lds_barrier();
        stage_out<L, CB>(ldd, (cpx<double> *)a.V[comp], cols, ncs);
        lds_barrier();
    }
}

// ---------------------------------------------------------------------------
// z pass of the RBC nonlinear terms (Compute_nlin_VxT, compute_diag_offdiag.cc
// :116-136, and the diag/offdiag products :47-112): the three velocity lines
// are loaded together, transformed once and held in registers for the six
// velocity products, then theta for the three u theta products; four c2r
// and nine r2c per line.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rbc_pair(int ia, int ib)
{
    // out: 0 xx, 1 yy, 2 zz, 3 xy, 4 yz, 5 zx, 6 x theta, 7 y theta, 8 z theta
    if (ia == ib) return ia;
    if (ia == 0) return ib == 1 ? 3 : (ib == 2 ? 5 : 6);
    if (ia == 1) return ib == 2 ? 4 : 7;
    return 8;
}

__device__ __forceinline__ void *sel_ptr(void *const *p, int i, int n)
{
    void *r = p[0];
    for (int k = 1; k < n; k++)
        if (i == k) r = p[k];
    return r;
}

// Two waves per SIMD also for FP64 512-point lines (a few spilled registers
// there; 19.3 ms per stage at 512x1024x1024 against 21.0 ms at one wave).
// SHIFT: an RK stage's velocity products are the trace-shifted u_x^2 - u_z^2,
// u_y^2 - u_z^2 and the off-diagonal ones (out[2], zz, not written): the stage
// uses the velocity term only projected, and the SFF projection removes the
// gradient of u_z^2 that the shift adds.
template <typename T, int M, bool SHIFT = false>
__global__ void __launch_bounds__(256, 2) k_z_rbc(ZArgs a)
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
    const size_t off = lsafe * a.g.nzp;
    cpx<T> u[3][8], xm[3];
#pragma unroll
    for (int c = 0; c < 3; c++) zline_load<T, M>(u[c], xm[c], (const cpx<T> *)a.in[c] + off, j);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        zline_c2r_from<T, M>(u[c], xm[c], L, twm, twn, j);
        line_sync<M / 8>();
    }
    auto emit = [&](const cpx<T> *x, const cpx<T> *y, int o) {
        cpx<T> p[8];
#pragma unroll
        for (int s = 0; s < 8; s++) p[s] = mk<T>(x[s].x * y[s].x, x[s].y * y[s].y);
        zline_r2c_post<T, M>(p, (cpx<T> *)a.out[o] + off, L, twm, twn, j, valid);
    };
    if constexpr (SHIFT) {
        // the off-diagonal products first, then the shifted squares (u_z is
        // last used there)
        emit(u[0], u[1], rbc_pair(0, 1));
        emit(u[1], u[2], rbc_pair(1, 2));
        emit(u[0], u[2], rbc_pair(0, 2));
#pragma unroll
        for (int ia = 0; ia < 2; ia++) {
            cpx<T> p[8];
#pragma unroll
            for (int s = 0; s < 8; s++)
                p[s] = mk<T>(u[ia][s].x * u[ia][s].x - u[2][s].x * u[2][s].x,
                             u[ia][s].y * u[ia][s].y - u[2][s].y * u[2][s].y);
            zline_r2c_post<T, M>(p, (cpx<T> *)a.out[ia] + off, L, twm, twn, j, valid);
        }
    } else {
#pragma unroll
        for (int ia = 0; ia < 3; ia++)
#pragma unroll
            for (int ib = ia; ib < 3; ib++) emit(u[ia], u[ib], rbc_pair(ia, ib));
    }
    cpx<T> t[8];
    zline_c2r<T, M>(t, (const cpx<T> *)a.in[3] + off, L, twm, twn, j);
    line_sync<M / 8>();
#pragma unroll
    for (int ia = 0; ia < 3; ia++) emit(u[ia], t, rbc_pair(ia, 3));
}

// ---------------------------------------------------------------------------
// Fused x pass of an RBC stage.
// ---------------------------------------------------------------------------
struct XrbcArgs {
    void *P[9];
    double *V[4];
    double *acc[4];
    void *U[4];
    const double *f_in[2];
    double *f_out[2];
    double *out[4];
    const void *tw, *twq;
    Grid g;
    RbcParams rp;
};

// Contribution of a transformed product q to one output component:
// sel 0 none, 1 +Kx q (d/dx of an 'S' product), 2 -Kx q ('C' product),
// 3 i Ky q, 4 i Kz q (SFF_pencil_tr.cc:77-151).  Branch-free form: the
// term is a q + i b q with a = ax Kx, b = ay Ky + az Kz and one of ax, ay,
// az equal to +-1 (the others 0), so every term keeps its exact value.
template <typename T>
struct RbcTerm {
    T ax, b;
    __device__ __forceinline__ RbcTerm(int sel, T Ky, T Kz)
        : ax(sel == 1 ? T(1) : (sel == 2 ? T(-1) : T(0))), b(sel == 3 ? Ky : (sel == 4 ? Kz : T(0)))
    {
    }
    __device__ __forceinline__ cpx<T> operator()(T Kx, cpx<T> q) const
    {
        const T a = ax * Kx;
        return mk<T>(a * q.x - b * q.y, a * q.y + b * q.x);
    }
};

// Selectors of the velocity products (xx C, yy C, zz C, xy S, yz C, zx S)
// for N_x, N_y, N_z in the reference's accumulation order.
__device__ __forceinline__ int rbc_sel(int p, int comp)
{
    switch (p) {
        case 0: return comp == 0 ? 2 : 0;                      // xx: N_x = -Kx F_C
        case 1: return comp == 1 ? 3 : 0;                      // yy: N_y = i Ky F_C
        case 2: return comp == 2 ? 4 : 0;                      // zz: N_z = i Kz F_C
        case 3: return comp == 0 ? 3 : (comp == 1 ? 1 : 0);    // xy: N_x += i Ky, N_y += Kx (F_S)
        case 4: return comp == 1 ? 4 : (comp == 2 ? 3 : 0);    // yz: N_y += i Kz, N_z += i Ky (F_C)
        default: return comp == 0 ? 4 : (comp == 2 ? 1 : 0);  // zx: N_x += i Kz, N_z += Kx (F_S)
    }
}

// One component of the Lawson stage (RK_SF.cc:97-215 with
// Tadvance_functions.cc:298-324, 566-570): r = right-hand side (FP64).
template <typename T, int L, int CB>
__device__ __forceinline__ void rbc_lawson(const Grid &g, const RbcParams &rp, int c, int j, int ly, int lz,
                                           double Ky, double Kz, const int *colw, const int *cols,
                                           cpx<double> *ldd, cpx<double> *ldb, cpx<T> *ldt, const cpx<T> *tw,
                                           QuarterTw<T> twq, const cpx<double> *r, double fh, double f1,
                                           double *V, double *acc, void *U, double *fo, bool sine)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    const int stage = rp.stage;
    const int ncw = g.ny * g.nzp, ncs = g.ny * g.nzc;
    const double dt = rp.dt;
    const double c_half = 0.5 * dt, c_one = 1.0 * dt;
    const double c_sixth = (1.0 / 6) * dt, c_third = (1.0 / 3) * dt;
    stage_in<L, CB>(ldd, (const cpx<double> *)V, cols, ncs);
    if (stage >= 2) stage_in<L, CB>(ldb, (const cpx<double> *)acc, colw, ncw);
    lds_barrier();
    cpx<T> v[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int lx = j + TPL * s;
        const int li = c * LP + lpad(lx);
        const double Kx = lx * g.kf[0];
        const double ks = (Kx * Kx + Ky * Ky) + Kz * Kz;
        const double Eh = exp(fh * ks);
        const cpx<double> v0 = ldd[li];
        cpx<double> vn, ac;
        if (stage == 1) {
            cpx<double> d0 = v0;
            if (dealias_zero(g, lx, ly, lz)) d0 = mk<double>(0, 0);
            const double E1 = exp(f1 * ks);
            ac = c_sixth * (E1 * r[s]);
            vn = Eh * (d0 + c_half * r[s]);
        } else if (stage == 2) {
            ac = ldb[li] + c_third * (Eh * r[s]);
            vn = Eh * v0 + c_half * r[s];
        } else if (stage == 3) {
            ac = ldb[li] + c_third * (Eh * r[s]);
            vn = Eh * (Eh * v0 + c_one * r[s]);
        } else {
            ac = ldb[li] + c_sixth * r[s];
            const double E1 = exp(f1 * ks);
            vn = E1 * v0 + ac;
        }
        // component-wise select (selecting the objects keeps them in memory)
        ldb[li] = mk<double>(stage == 4 ? vn.x : ac.x, stage == 4 ? vn.y : ac.y);
        if (stage == 4 && !(isfinite(vn.x) && isfinite(vn.y)) && rp.nonfinite) *rp.nonfinite = 1;
        // the undealiased next-stage value (this field's share of the next
        // force) replaces v0 in the thread's own slot of the state tile
        if (fo) ldd[li] = vn;
        if (dealias_zero(g, lx, ly, lz)) vn = mk<double>(0, 0);
        v[s] = cvt<T>(vn);
    }
    lds_barrier();
    if (stage == 4) stage_out<L, CB>(ldb, (cpx<double> *)V, cols, ncs);
    else stage_out<L, CB>(ldb, (cpx<double> *)acc, colw, ncw);
    if (fo) {
        stage_out<L, CB>(ldd, (cpx<double> *)fo, cols, ncs);
        lds_barrier();
    }
    sc_inverse<L>(v, ldt + c * LP, tw, twq, j, sine);
    lds_barrier();
    stage_out<L, CB>(ldt, (cpx<T> *)U, colw, ncw);
    lds_barrier();
}

// PHASE 0: temperature (N_T, its right-hand side, Lawson for theta);
// PHASE 1: velocity.  The two launches of a stage touch disjoint data (the
// force inputs they read are the other field's, of this stage's input).
template <typename T, int L, int CB, int PHASE>
__global__ void __launch_bounds__(CB *(L / 8), 2) k_x_rbc(XrbcArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    // A: FFT exchanges and state staging; B: products arriving by
    // direct-to-LDS loads, then the accumulator tile (separate LDS objects)
    __shared__ __attribute__((aligned(16))) char rawA[CB * LP * sizeof(cpx<double>)];
    __shared__ __attribute__((aligned(16))) char rawB[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    __shared__ cpx<T> twl[L], twqlo[L / 8], twqhi[8];
    cpx<double> *ldd = (cpx<double> *)rawA;
    cpx<T> *ldt = (cpx<T> *)rawA;
    cpx<double> *ldb = (cpx<double> *)rawB;
    constexpr bool DMA = tile_dma_ok<T, L, CB>();
    const Grid &g = a.g;
    const RbcParams &rp = a.rp;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= x_tiles(g, CB, false)) return;
    tile_columns<CB>(colw, cols, g, false, tile);
    for (int i = threadIdx.x; i < L; i += CB * TPL) twl[i] = ((const cpx<T> *)a.tw)[i];
    if (threadIdx.x < TPL) twqlo[threadIdx.x] = ((const cpx<T> *)a.twq)[threadIdx.x];
    if (threadIdx.x < 8) twqhi[threadIdx.x] = ((const cpx<T> *)a.twq)[TPL * threadIdx.x];
    const QuarterTw<T> twql{twqlo, twqhi};
    lds_barrier();
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int cw = colw[c];
    const int ly = cw >= 0 ? cw / g.nzp : 0, lz = cw >= 0 ? cw % g.nzp : 0;
    const double Ky = kwave(ly, g.ny) * g.kf[1], Kz = lz * g.kf[2];
    const T Kyt = (T)Ky, Kzt = (T)Kz;
    const T sc = (T)rp.scale;
    const int ncw = g.ny * g.nzp, ncs = g.ny * g.nzc;
    const int stage = rp.stage;
    const bool want_rhs = stage > 0 || rp.write_rhs;
    const int cw0 = tile * CB;
    // Product p of a phase into registers v (Makhoul order, sign-alternated
    // for sine products): from B after its direct-to-LDS load (then the load
    // of product p_next starts), or staged through A.
    auto load_product = [&](cpx<T> *v, int p, int p_next, int jj, bool sine) {
        if constexpr (DMA) {
            dma_wait();
            lds_barrier();
            sc_gather<L, T>(v, jj, sine, [&](int r) { return dma_tile_read<T, L, CB>(rawB, r, c); });
            lds_barrier();
            if (p_next >= 0) dma_tile<T, L, CB>((const cpx<T> *)sel_ptr(a.P, p_next, 9), cw0, ncw, rawB);
        } else {
            stage_in<L, CB>(ldt, (const cpx<T> *)sel_ptr(a.P, p, 9), colw, ncw);
            lds_barrier();
            sc_gather<L, T>(v, jj, sine, [&](int r) { return ldt[c * LP + lpad(r)]; });
            lds_barrier();
        }
    };

    // ===== temperature: N_T = -Kx F_C(u_x T) + i Ky F_S(u_y T) + i Kz F_S(u_z T) =====
    if constexpr (PHASE == 0) {
        cpx<T> nt[8];
#pragma unroll
        for (int s = 0; s < 8; s++) nt[s] = mk<T>(0, 0);
        if constexpr (DMA) dma_tile<T, L, CB>((const cpx<T> *)a.P[6], cw0, ncw, rawB);
#pragma unroll 1
        for (int p = 6; p < 9; p++) {
            int jj = j;  // hidden from loop-invariant code motion (see k_x_nlin)
            asm volatile("" : "+v"(jj));
            cpx<T> v[8];
            load_product(v, p, p < 8 ? p + 1 : -1, jj, p != 6);
            sc_forward_core<L>(v, ldt + c * LP, twl, twql, jj, p != 6);
            const RbcTerm<T> tt(p == 6 ? 2 : (p == 7 ? 3 : 4), Kyt, Kzt);
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const T Kx = (T)((jj + TPL * s) * g.kf[0]);
                nt[s] = nt[s] + tt(Kx, sc * v[s]);
            }
            if constexpr (!DMA) lds_barrier();
        }
        if constexpr (DMA) lds_barrier();
        cpx<double> r[8];
        if (want_rhs) {
            // RHS_T = -(N_T - F_T), F_T = u_x of the stage input
            stage_in<L, CB>(ldb, (const cpx<double> *)a.f_in[1], cols, ncs);
            lds_barrier();
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const cpx<double> n = cvt<double>(nt[s]) - ldb[c * LP + lpad(j + TPL * s)];
                r[s] = mk<double>(-n.x, -n.y);
            }
            lds_barrier();
        } else {
#pragma unroll
            for (int s = 0; s < 8; s++) r[s] = cvt<double>(nt[s]);
        }
        if (stage == 0) {
#pragma unroll
            for (int s = 0; s < 8; s++) ldd[c * LP + lpad(j + TPL * s)] = r[s];
            lds_barrier();
            stage_out<L, CB>(ldd, (cpx<double> *)a.out[3], cols, ncs);
            lds_barrier();
        } else {
            rbc_lawson<T, L, CB>(g, rp, c, j, ly, lz, Ky, Kz, colw, cols, ldd, ldb, ldt, twl, twql, r, rp.fh_t, rp.f1_t,
                                 a.V[3], a.acc[3], a.U[3], stage < 4 ? a.f_out[0] : nullptr, true);
        }
    }

    if constexpr (PHASE == 0) return;
    // ===== velocity: divergence-form N with the SFF parities =====
    cpx<T> N[3][8];
#pragma unroll
    for (int s = 0; s < 8; s++) N[0][s] = N[1][s] = N[2][s] = mk<T>(0, 0);
    if constexpr (DMA) dma_tile<T, L, CB>((const cpx<T> *)a.P[0], cw0, ncw, rawB);
    // trace-shifted products (an RK stage): xx - zz, yy - zz, xy, yz, zx, the
    // zz slot skipped
    const bool shift = rp.trace_free != 0;
    const int nvp = shift ? 5 : 6;
#pragma unroll 1
    for (int q = 0; q < nvp; q++) {
        const int p = shift && q >= 2 ? q + 1 : q;
        const int pn = q + 1 < nvp ? (shift && q + 1 >= 2 ? q + 2 : q + 1) : -1;
        int jj = j;
        asm volatile("" : "+v"(jj));
        cpx<T> v[8];
        load_product(v, p, pn, jj, p == 3 || p == 5);
        sc_forward_core<L>(v, ldt + c * LP, twl, twql, jj, p == 3 || p == 5);
        const RbcTerm<T> t0(rbc_sel(p, 0), Kyt, Kzt), t1(rbc_sel(p, 1), Kyt, Kzt), t2(rbc_sel(p, 2), Kyt, Kzt);
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const T Kx = (T)((jj + TPL * s) * g.kf[0]);
            const cpx<T> q = sc * v[s];
            N[0][s] = N[0][s] + t0(Kx, q);
            N[1][s] = N[1][s] + t1(Kx, q);
            N[2][s] = N[2][s] + t2(Kx, q);
        }
        if constexpr (!DMA) lds_barrier();
    }
    if constexpr (DMA) lds_barrier();
    if (want_rhs) {
        // nlin -= F (F_x = theta of the stage input); p = div/K^2 with the SFF
        // divergence (+Kx on the 'S' x component); nlin += grad p ('C' parity:
        // -Kx p); negate (Compute_rhs_scalar)
        stage_in<L, CB>(ldd, (const cpx<double> *)a.f_in[0], cols, ncs);
        lds_barrier();
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int lx = j + TPL * s;
            const double Kx = lx * g.kf[0];
            cpx<double> n0 = cvt<double>(N[0][s]) - ldd[c * LP + lpad(lx)];
            cpx<double> n1 = cvt<double>(N[1][s]), n2 = cvt<double>(N[2][s]);
            cpx<double> div = Kx * n0;
            div = div + mul_i(Ky, n1);
            div = div + mul_i(Kz, n2);
            const double ks = (Kx * Kx + Ky * Ky) + Kz * Kz;
            cpx<double> pr = mk<double>(0, 0);
            if (!(lx == 0 && ly == 0 && lz == 0)) pr = mk<double>(div.x / ks, div.y / ks);
            n0 = n0 + (-Kx) * pr;
            n1 = n1 + mul_i(Ky, pr);
            n2 = n2 + mul_i(Kz, pr);
            N[0][s] = cvt<T>(mk<double>(-n0.x, -n0.y));
            N[1][s] = cvt<T>(mk<double>(-n1.x, -n1.y));
            N[2][s] = cvt<T>(mk<double>(-n2.x, -n2.y));
        }
        lds_barrier();
    }
    // the three components unrolled: N[K] and the field pointers are static
    // (a runtime component index put N in scratch memory)
    auto vel_comp = [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        cpx<double> r[8];
#pragma unroll
        for (int s = 0; s < 8; s++) r[s] = cvt<double>(N[K][s]);
        if (stage == 0) {
#pragma unroll
            for (int s = 0; s < 8; s++) ldd[c * LP + lpad(j + TPL * s)] = r[s];
            lds_barrier();
            stage_out<L, CB>(ldd, (cpx<double> *)a.out[K], cols, ncs);
            lds_barrier();
        } else {
            int jj = j;  // keeps the per-thread LDS addresses inside each component
            asm volatile("" : "+v"(jj));
            rbc_lawson<T, L, CB>(g, rp, c, jj, ly, lz, Ky, Kz, colw, cols, ldd, ldb, ldt, twl, twql, r, rp.fh_u,
                                 rp.f1_u, a.V[K], a.acc[K], a.U[K], (K == 0 && stage < 4) ? a.f_out[1] : nullptr,
                                 K == 0);
        }
    };
    vel_comp(std::integral_constant<int, 0>{});
    vel_comp(std::integral_constant<int, 1>{});
    vel_comp(std::integral_constant<int, 2>{});
}

// ---------------------------------------------------------------------------
// Reality conditions on SFF fields
// ---------------------------------------------------------------------------
struct SffPlaneArgs {
    double *V[4];
    Grid g;
    int ncomp, mode;
};

#if TG_LEN_UNIT == 0
__global__ void k_sff_reality(SffPlaneArgs a)
{
    const Grid &g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nx * g.ny) return;
    const int lx = i / g.ny, ly = i % g.ny;
    const int ky = kwave(ly, g.ny);
    for (int comp = 0; comp < a.ncomp; comp++) {
        cpx<double> *A = (cpx<double> *)a.V[comp];
        const size_t row = ((size_t)lx * g.ny + ly) * g.nzc;
        // A(lx, ky, 0) = conj A(lx, -ky, 0) for -Ny/2 < ky < 0 reads a
        // ky > 0 row, which this kernel never writes; ky = 0 conjugates in place
        if (a.mode == 1 && ky <= 0) {
            const cpx<double> src = A[((size_t)lx * g.ny + (-ky)) * g.nzc];
            A[row] = conj(src);
        }
        A[row + g.nzc - 1] = mk<double>(0, 0);
    }
}

#endif  // TG_LEN_UNIT == 0

// ---------------------------------------------------------------------------
// Per-length launch functions and launchers
// ---------------------------------------------------------------------------
template <typename T, int LL>
hipError_t xinv_sc_len(const XscArgs &a, hipStream_t st)
{
    constexpr int CB = col_block<T>(LL, sizeof(cpx<double>));
    hipLaunchKernelGGL((k_xinv_sc<T, LL, CB>), dim3(x_grid(x_tiles(a.g, CB, (a.flags & 4) != 0))),
                       dim3(CB * (LL / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int LL>
hipError_t xfwd_sc_len(const XscArgs &a, hipStream_t st)
{
    constexpr int CB = col_block<T>(LL, sizeof(cpx<double>));
    hipLaunchKernelGGL((k_xfwd_sc<T, LL, CB>), dim3(x_grid(x_tiles(a.g, CB, false))), dim3(CB * (LL / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int MM>
hipError_t z_rbc_len(const ZArgs &a, hipStream_t st)
{
    constexpr int LPB = z_lines_per_block<MM>();
    const size_t nlines = (size_t)a.g.nx * a.g.ny;
    dim3 grid((unsigned)((nlines + LPB - 1) / LPB));
    if (a.ncomp == 8) hipLaunchKernelGGL((k_z_rbc<T, MM, true>), grid, dim3(LPB * (MM / 8)), 0, st, a);
    else hipLaunchKernelGGL((k_z_rbc<T, MM>), grid, dim3(LPB * (MM / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int LL>
hipError_t x_rbc_len(const XrbcArgs &a, hipStream_t st)
{
    constexpr int CB = col_block<T>(LL, 2 * sizeof(cpx<double>));
    const dim3 grid(x_grid(x_tiles(a.g, CB, false)));
    hipLaunchKernelGGL((k_x_rbc<T, LL, CB, 0>), grid, dim3(CB * (LL / 8)), 0, st, a);
    hipLaunchKernelGGL((k_x_rbc<T, LL, CB, 1>), grid, dim3(CB * (LL / 8)), 0, st, a);
    return hipGetLastError();
}

#define TG_SFF_L(N, T, P)                                                   \
    P template hipError_t xinv_sc_len<T, N>(const XscArgs &, hipStream_t); \
    P template hipError_t xfwd_sc_len<T, N>(const XscArgs &, hipStream_t); \
    P template hipError_t x_rbc_len<T, N>(const XrbcArgs &, hipStream_t);
#define TG_SFF_M(N, T, P) P template hipError_t z_rbc_len<T, N>(const ZArgs &, hipStream_t);

#if TG_LEN_UNIT == 0
TG_ALL_UNITS(TG_SFF_L, TG_SFF_M, extern)

template <typename T>
hipError_t launch_xinv_sc(const Grid &g, double *const *V, int ncomp, int parity, void *const *U, int flags,
                          const void *twx, const void *twq, hipStream_t st)
{
    XscArgs a{};
    for (int i = 0; i < ncomp; i++) {
        a.V[i] = V[i];
        a.U[i] = U[i];
    }
    a.tw = twx;
    a.twq = twq;
    a.g = g;
    a.ncomp = ncomp;
    a.parity = parity;
    a.flags = flags;
    TG_DISPATCH_L(T, g.nx, return (xinv_sc_len<T, LL>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_xfwd_sc(const Grid &g, void *const *W, int ncomp, int parity, double *const *out, double scale,
                          const void *twx, const void *twq, hipStream_t st)
{
    XscArgs a{};
    for (int i = 0; i < ncomp; i++) {
        a.U[i] = W[i];
        a.V[i] = out[i];
    }
    a.tw = twx;
    a.twq = twq;
    a.g = g;
    a.ncomp = ncomp;
    a.parity = parity;
    a.scale = scale;
    TG_DISPATCH_L(T, g.nx, return (xfwd_sc_len<T, LL>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_z_rbc(const Grid &g, void *const *U, void *const *P, const void *twm, const void *twn,
                        hipStream_t st, bool shift)
{
    ZArgs a{};
    a.ncomp = shift ? 8 : 9;
    for (int i = 0; i < 4; i++) a.in[i] = U[i];
    for (int i = 0; i < 9; i++) a.out[i] = P[i];
    a.twm = twm;
    a.twn = twn;
    a.g = g;
    TG_DISPATCH_M(T, g.nz / 2, return (z_rbc_len<T, MM>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_x_rbc(const Grid &g, const RbcParams &rp, void *const *P, double *const *V, double *const *acc,
                        void *const *U, const double *const *f_in, double *const *f_out, double *const *out,
                        const void *twx, const void *twq, hipStream_t st)
{
    XrbcArgs a{};
    for (int i = 0; i < 9; i++) a.P[i] = P[i];
    for (int i = 0; i < 4; i++) {
        a.V[i] = V ? V[i] : nullptr;
        a.acc[i] = acc ? acc[i] : nullptr;
        a.U[i] = U ? U[i] : nullptr;
        a.out[i] = out ? out[i] : nullptr;
    }
    for (int i = 0; i < 2; i++) {
        a.f_in[i] = f_in ? f_in[i] : nullptr;
        a.f_out[i] = f_out ? f_out[i] : nullptr;
    }
    a.tw = twx;
    a.twq = twq;
    a.g = g;
    a.rp = rp;
    TG_DISPATCH_L(T, g.nx, return (x_rbc_len<T, LL>)(a, st));
    return hipErrorInvalidValue;
}

hipError_t launch_sff_reality(const Grid &g, double *const *V, int ncomp, int mode, hipStream_t st)
{
    SffPlaneArgs a{};
    for (int i = 0; i < ncomp; i++) a.V[i] = V[i];
    a.g = g;
    a.ncomp = ncomp;
    a.mode = mode;
    const int n = g.nx * g.ny;
    hipLaunchKernelGGL(k_sff_reality, dim3((n + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

#define TG_SFF_INST(T)                                                                                              \
    template hipError_t launch_xinv_sc<T>(const Grid &, double *const *, int, int, void *const *, int, const void *, \
                                          const void *, hipStream_t);                                               \
    template hipError_t launch_xfwd_sc<T>(const Grid &, void *const *, int, int, double *const *, double,           \
                                          const void *, const void *, hipStream_t);                                 \
    template hipError_t launch_z_rbc<T>(const Grid &, void *const *, void *const *, const void *, const void *,     \
                                        hipStream_t, bool);                                                         \
    template hipError_t launch_x_rbc<T>(const Grid &, const RbcParams &, void *const *, double *const *,            \
                                        double *const *, void *const *, const double *const *, double *const *,     \
                                        double *const *, const void *, const void *, hipStream_t);

TG_SFF_INST(double)
TG_SFF_INST(float)
#else
TG_THIS_UNIT(TG_SFF_L, TG_SFF_M)
#endif  // TG_LEN_UNIT

}  // namespace tg
