// tg_adjoint.hip -- adjoint of the NS right-hand side f(V) = -P NL(D V),
// NL_i = sum_j i K_j F[u_j u_i], u = I(D V), and of the Lawson RK4 step.
// (No counterpart in the reference.)
//
// Inner product: the real dot product of the stored half-spectrum (Re, Im).
// With c = 1 on the kz = 0 and kz = Nz/2 planes and 2 elsewhere the
// transposes of the transforms are I^T = N^3 diag(c) F and
// F^T = N^-3 I diag(1/c); (i K)^T = -i K; P and D are symmetric.  So
//   f'(V)^T G = -D diag(c) F[q],  q_m = sum_j u_j sigma_jm,
//   sigma_jm = I(diag(1/c)(-i K_m M_j - i K_j M_m)),  M = P G,
// which needs the stage's u (taped by the forward sweep), five inverse and
// three forward transforms: x-inverse of the symmetrised fields (built from
// M in registers), y-inverse, z c2r + products with the tape + r2c,
// y-forward, x-forward with the stage epilogue.  sigma is trace-free
// (sigma_jj summed is -2i K.M/c = 0, M being projected), so sigma_zz =
// -sigma_xx - sigma_yy is not transformed: the z pass subtracts u_z sigma_xx
// and u_z sigma_yy from q_z instead.
#include "tg_adjoint.h"
#include "tg_device.h"
#include "tg_pass.h"

namespace tg {

struct XsymArgs {
    const double *G[3];
    void *S[5];
    const void *tw;
    Grid g;
};

// From the cotangent G of a tile (put(k) stages component k of G into the
// LDS tile in the stage_in layout): M = P G, the five symmetrised fields
// xx, yy, xy, yz, zx of (-i K_m M_j - i K_j M_m)/c (zz = -xx - yy), their
// x-inverse, stored into S.
template <typename T, int L, int CB, typename Put>
__device__ __forceinline__ void xsym_tile(const Grid &g, char *raw, const int *colw, const cpx<T> *tw,
                                          void *const *S, Put put)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    cpx<double> *ldd = (cpx<double> *)raw;
    cpx<T> *ldt = (cpx<T> *)raw;
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int cw = colw[c];
    const int ly = cw >= 0 ? cw / g.nzp : 0, lz = cw >= 0 ? cw % g.nzp : 0;
    const double Ky = kwave(ly, g.ny) * g.kf[1], Kz = lz * g.kf[2];
    const int ncw = g.ny * g.nzp;
    cpx<double> m[3][8];
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        put(comp);
        lds_barrier();
#pragma unroll
        for (int s = 0; s < 8; s++) m[comp][s] = ldd[c * LP + lpad(j + TPL * s)];
        lds_barrier();
    }
    // M = P G (Pressure.cc:56-72 form: G - K (K.G)/K^2, identity at K = 0)
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const double Kx = kwave(j + TPL * s, g.nx) * g.kf[0];
        const double ks = (Kx * Kx + Ky * Ky) + Kz * Kz;
        if (ks > 0.0) {
            cpx<double> d = Kx * m[0][s];
            d = d + Ky * m[1][s];
            d = d + Kz * m[2][s];
            const cpx<double> e = mk<double>(d.x / ks, d.y / ks);
            m[0][s] = m[0][s] - Kx * e;
            m[1][s] = m[1][s] - Ky * e;
            m[2][s] = m[2][s] - Kz * e;
        }
    }
    const double w = (lz == 0 || lz == g.nz / 2) ? 1.0 : 0.5;
    constexpr int pa[5] = {0, 1, 0, 1, 2};
    constexpr int pb[5] = {0, 1, 1, 2, 0};
#pragma unroll
    for (int t = 0; t < 5; t++) {
        cpx<T> v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const double Kx = kwave(j + TPL * s, g.nx) * g.kf[0];
            const double K[3] = {Kx, Ky, Kz};
            const cpx<double> A = m[pa[t]][s], B = m[pb[t]][s];
            // -i K_b A - i K_a B, times 1/c
            const double re = K[pb[t]] * A.y + K[pa[t]] * B.y;
            const double im = -(K[pb[t]] * A.x + K[pa[t]] * B.x);
            v[s] = (cw >= 0) ? mk<T>((T)(w * re), (T)(w * im)) : mk<T>(0, 0);
        }
        fft_line<L, +1>(v, ldt + c * LP, tw, j);
#pragma unroll
        for (int s = 0; s < 8; s++) ldt[c * LP + lpad(j + TPL * s)] = v[s];
        lds_barrier();
        stage_out<L, CB>(ldt, (cpx<T> *)S[t], colw, ncw);
        lds_barrier();
    }
}

template <typename T, int L, int CB>
__global__ void __launch_bounds__(CB *(L / 8)) k_xinv_sym(XsymArgs a)
{
    constexpr int LP = lpad_len(L);
    __shared__ __attribute__((aligned(16))) char raw[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    const Grid &g = a.g;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= x_tiles(g, CB, false)) return;
    tile_columns<CB>(colw, cols, g, false, tile);
    lds_barrier();
    const int ncs = g.ny * g.nzc;
    // the three cotangent tiles' loads are in flight together
    cpx<double> gr[3][8];
#pragma unroll
    for (int comp = 0; comp < 3; comp++) stage_load<L, CB>(gr[comp], (const cpx<double> *)a.G[comp], cols, ncs);
    xsym_tile<T, L, CB>(g, raw, colw, (const cpx<T> *)a.tw, a.S,
                        [&](int k) { stage_put<L, CB>((cpx<double> *)raw, gr[k], cols); });
}

// q += u * v (SUB: q -= u * v) on the packed real points of this thread (u
// from the tape line).
template <typename T, int M, bool SUB = false>
__device__ __forceinline__ void tape_mac(cpx<T> *q, const cpx<T> *v, const cpx<T> *__restrict__ u, int j)
{
    constexpr int TPL = M / 8;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const cpx<T> A = u[j + TPL * s];
        const cpx<T> d = mk<T>(A.x * v[s].x, A.y * v[s].y);
        q[s] = SUB ? q[s] - d : q[s] + d;
    }
}

// sigma lines in a.in[]: 0 xx, 1 yy, 2 xy, 3 yz, 4 zx (zz = -xx - yy).  The
// three accumulators q_m = sum_j u_j sigma_jm are live together, so each
// sigma line takes one c2r: a diagonal line sigma_jj adds u_j sigma_jj to q_j
// and, as its share of sigma_zz, subtracts u_z sigma_jj from q_z; an
// off-diagonal line sigma_jm adds u_j sigma_jm to q_m and u_m sigma_jm to
// q_j.  Five c2r and three r2c per line.
template <typename T, int M>
__global__ void __launch_bounds__(256) k_z_adj(ZArgs a)
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
    cpx<T> q[3][8];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int s = 0; s < 8; s++) q[m][s] = mk<T>(0, 0);
#pragma unroll 1
    for (int t = 0; t < 5; t++) {
        // sigma_jm of line t: (j, m) = (t, t) for t < 2, (t-2, t-1 mod 3) after
        const int jt = t < 2 ? t : t - 2, mt = t < 2 ? t : (t == 4 ? 0 : t - 1);
        const cpx<T> *X = (const cpx<T> *)(t == 0 ? a.in[0] : t == 1 ? a.in[1] : t == 2 ? a.in[2] :
                                            t == 3 ? a.in[3] : a.in[4]);
        cpx<T> v[8];
        zline_c2r<T, M>(v, X + off, L, twm, twn, j);
        line_sync<M / 8>();
        const cpx<T> *uj = (const cpx<T> *)(jt == 0 ? a.tape[0] : jt == 1 ? a.tape[1] : a.tape[2]) + lsafe * M;
        const cpx<T> *um = (const cpx<T> *)(mt == 0 ? a.tape[0] : mt == 1 ? a.tape[1] : a.tape[2]) + lsafe * M;
        const cpx<T> *uz = (const cpx<T> *)a.tape[2] + lsafe * M;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (k == mt) tape_mac<T, M>(q[k], v, uj, j);
            if (t >= 2 && k == jt) tape_mac<T, M>(q[k], v, um, j);
        }
        if (t < 2) tape_mac<T, M, true>(q[2], v, uz, j);
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
        cpx<T> *Q = (cpx<T> *)(m == 0 ? a.out[0] : m == 1 ? a.out[1] : a.out[2]);
        zline_r2c_post<T, M>(q[m], Q + off, L, twm, twn, j, valid);
    }
}

struct XadjArgs {
    void *Q[3];
    double *lam[3], *acc[3], *lg[3];
    void *S[5];  // FUSE: the next stage's five symmetrised fields
    const void *tw;
    Grid g;
    AdjParams ap;
};

// FUSE: the next stage's cotangent (lg, or lam after the step's last stage)
// stays in registers instead of going through HBM, and the next stage's
// xsym_tile runs on it here (lg is not stored).
template <typename T, int L, int CB, bool FUSE>
__global__ void __launch_bounds__(CB *(L / 8)) k_xfwd_adj(XadjArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    __shared__ __attribute__((aligned(16))) char raw[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    cpx<double> *ldd = (cpx<double> *)raw;
    cpx<T> *ldt = (cpx<T> *)raw;
    const Grid &g = a.g;
    const AdjParams &ap = a.ap;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= x_tiles(g, CB, false)) return;
    tile_columns<CB>(colw, cols, g, false, tile);
    lds_barrier();
    const int c = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const cpx<T> *tw = (const cpx<T> *)a.tw;
    const int ncw = g.ny * g.nzp, ncs = g.ny * g.nzc;
    const double c_sixth = (1.0 / 6) * ap.dt, c_third = (1.0 / 3) * ap.dt, c_half = 0.5 * ap.dt;
    cpx<double> gn[3][8];  // FUSE: next cotangent, stage_in layout
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
#pragma unroll
        for (int ii = 0; ii < 8; ii++) gn[comp][ii] = mk<double>(0, 0);
        stage_in<L, CB>(ldt, (const cpx<T> *)a.Q[comp], colw, ncw);
        lds_barrier();
        cpx<T> v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) v[s] = ldt[c * LP + lpad(j + TPL * s)];
        lds_barrier();
        fft_line<L, -1>(v, ldt + c * LP, tw, j);
        lds_barrier();
#pragma unroll
        for (int s = 0; s < 8; s++) ldd[c * LP + lpad(j + TPL * s)] = ap.scale * cvt<double>(v[s]);
        lds_barrier();
        cpx<double> *lam = (cpx<double> *)a.lam[comp], *acc = (cpx<double> *)a.acc[comp];
        cpx<double> *lg = (cpx<double> *)a.lg[comp];
        // thread t: column t % CB, rows t / CB + TPL i; the eight elements'
        // inputs are loaded before the first result is stored
        const int cc = threadIdx.x % CB, r0 = threadIdx.x / CB, col = cols[cc];
        if (col >= 0) {
            const int ly = col / g.nzc, lz = col % g.nzc;
            const double cw = (lz == 0 || lz == g.nz / 2) ? -1.0 : -2.0;
            const double Ky = kwave(ly, g.ny) * g.kf[1], Kz = lz * g.kf[2];
            const bool need_l = ap.mode != ADJ_RHS;
            const bool need_a = ap.mode == ADJ_S2 || ap.mode == ADJ_S1 || ap.mode == ADJ_S0;
            cpx<double> l4[8], a4[8];
#pragma unroll
            for (int ii = 0; ii < 8; ii++) {
                const size_t i = (size_t)(r0 + TPL * ii) * ncs + col;
                l4[ii] = need_l ? mk<double>(lam[i].x, lam[i].y) : mk<double>(0, 0);
                a4[ii] = need_a ? mk<double>(acc[i].x, acc[i].y) : mk<double>(0, 0);
            }
#pragma unroll
            for (int ii = 0; ii < 8; ii++) {
                const int lx = r0 + TPL * ii;
                const size_t i = (size_t)lx * ncs + col;
                cpx<double> r = ldd[cc * LP + lpad(lx)];
                r = dealias_zero(g, lx, ly, lz) ? mk<double>(0, 0) : cw * r;
                if (ap.mode == ADJ_RHS) {
                    lg[i] = r;
                    continue;
                }
                const double Kx = kwave(lx, g.nx) * g.kf[0];
                const double ks = (Kx * Kx + Ky * Ky) + Kz * Kz;
                cpx<double> nx;
                if (ap.mode == ADJ_S3) {
                    const double Eh = visc_exp(ap.fh, ap.fhh, ap.hq, ks);
                    const cpx<double> lv = c_sixth * r;
                    acc[i] = (Eh * Eh) * lv;
                    nx = c_third * (Eh * l4[ii]) + ap.dt * (Eh * lv);
                } else if (ap.mode == ADJ_S2) {
                    const double Eh = visc_exp(ap.fh, ap.fhh, ap.hq, ks);
                    acc[i] = a4[ii] + Eh * r;
                    nx = c_third * (Eh * l4[ii]) + c_half * r;
                } else if (ap.mode == ADJ_S1) {
                    const double Eh = visc_exp(ap.fh, ap.fhh, ap.hq, ks), E1 = visc_exp(ap.f1, ap.f1h, ap.hq, ks);
                    acc[i] = a4[ii] + Eh * r;
                    nx = c_sixth * (E1 * l4[ii]) + c_half * (Eh * r);
                } else {
                    const double E1 = visc_exp(ap.f1, ap.f1h, ap.hq, ks);
                    nx = (E1 * l4[ii] + a4[ii]) + r;
                    lam[i] = nx;
                }
                if (ap.mode != ADJ_S0 && !FUSE) lg[i] = nx;
                if constexpr (FUSE) gn[comp][ii] = nx;
            }
        }
        lds_barrier();
    }
    if constexpr (FUSE)
        xsym_tile<T, L, CB>(g, raw, colw, tw, a.S, [&](int k) { stage_put<L, CB>(ldd, gn[k], cols); });
}

template <typename T, int LL>
hipError_t xinv_sym_len(const XsymArgs &a, hipStream_t st)
{
    constexpr int CB = col_block<T>(LL, sizeof(cpx<double>));
    hipLaunchKernelGGL((k_xinv_sym<T, LL, CB>), dim3(x_grid(x_tiles(a.g, CB, false))), dim3(CB * (LL / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int MM>
hipError_t z_adj_len(const ZArgs &a, hipStream_t st)
{
    constexpr int LPB = z_lines_per_block<MM>();
    const size_t nlines = (size_t)a.g.nx * a.g.ny;
    dim3 grid((unsigned)((nlines + LPB - 1) / LPB));
    hipLaunchKernelGGL((k_z_adj<T, MM>), grid, dim3(LPB * (MM / 8)), 0, st, a);
    return hipGetLastError();
}

template <typename T, int LL>
hipError_t xfwd_adj_len(const XadjArgs &a, bool fuse, hipStream_t st)
{
    constexpr int CB = col_block<T>(LL, sizeof(cpx<double>));
    const dim3 grid(x_grid(x_tiles(a.g, CB, false))), block(CB * (LL / 8));
    if (fuse) hipLaunchKernelGGL((k_xfwd_adj<T, LL, CB, true>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_xfwd_adj<T, LL, CB, false>), grid, block, 0, st, a);
    return hipGetLastError();
}

#define TG_ADJ_L(N, T, P)                                                     \
    P template hipError_t xinv_sym_len<T, N>(const XsymArgs &, hipStream_t); \
    P template hipError_t xfwd_adj_len<T, N>(const XadjArgs &, bool, hipStream_t);
#define TG_ADJ_M(N, T, P) P template hipError_t z_adj_len<T, N>(const ZArgs &, hipStream_t);

#if TG_LEN_UNIT == 0
TG_ALL_UNITS(TG_ADJ_L, TG_ADJ_M, extern)

template <typename T>
hipError_t launch_xinv_sym(const Grid &g, double *const *G, void *const *S, const void *twx, hipStream_t st)
{
    XsymArgs a{};
    for (int i = 0; i < 3; i++) a.G[i] = G[i];
    for (int i = 0; i < 5; i++) a.S[i] = S[i];
    a.tw = twx;
    a.g = g;
    TG_DISPATCH_L(T, g.nx, return (xinv_sym_len<T, LL>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_z_adj(const Grid &g, void *const *S, void *const *Q, void *const *tape, const void *twm,
                        const void *twn, hipStream_t st)
{
    ZArgs a{};
    for (int i = 0; i < 5; i++) a.in[i] = S[i];
    for (int i = 0; i < 3; i++) {
        a.out[i] = Q[i];
        a.tape[i] = tape[i];
    }
    a.twm = twm;
    a.twn = twn;
    a.g = g;
    TG_DISPATCH_M(T, g.nz / 2, return (z_adj_len<T, MM>)(a, st));
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_xfwd_adj(const Grid &g, const AdjParams &ap, void *const *Q, double *const *lam,
                           double *const *acc, double *const *lg, const void *twx, hipStream_t st,
                           void *const *S_next)
{
    XadjArgs a{};
    if (S_next && ap.mode == ADJ_RHS) return hipErrorInvalidValue;
    for (int i = 0; i < 5; i++) a.S[i] = S_next ? S_next[i] : nullptr;
    for (int i = 0; i < 3; i++) {
        a.Q[i] = Q[i];
        a.lam[i] = lam ? lam[i] : nullptr;
        a.acc[i] = acc ? acc[i] : nullptr;
        a.lg[i] = lg ? lg[i] : nullptr;
    }
    a.tw = twx;
    a.g = g;
    a.ap = ap;
    TG_DISPATCH_L(T, g.nx, return (xfwd_adj_len<T, LL>)(a, S_next != nullptr, st));
    return hipErrorInvalidValue;
}

#define TG_ADJ_INST(T)                                                                                           \
    template hipError_t launch_xinv_sym<T>(const Grid &, double *const *, void *const *, const void *, hipStream_t); \
    template hipError_t launch_z_adj<T>(const Grid &, void *const *, void *const *, void *const *, const void *,  \
                                        const void *, hipStream_t);                                              \
    template hipError_t launch_xfwd_adj<T>(const Grid &, const AdjParams &, void *const *, double *const *,       \
                                           double *const *, double *const *, const void *, hipStream_t,         \
                                           void *const *);

TG_ADJ_INST(double)
TG_ADJ_INST(float)
#else
TG_THIS_UNIT(TG_ADJ_L, TG_ADJ_M)
#endif  // TG_LEN_UNIT

}  // namespace tg
