// tg_pass.h -- device helpers shared by the x, y and z passes (tiling of the
// x pass through LDS, z-line real<->complex transforms, wavenumbers and the
// dealias mask) and the size dispatch of the launchers.
#pragma once
#include "tg_device.h"
#include "tg_kernels.h"

namespace tg {

static constexpr int kLdsBudget = 73728;  // bytes; two blocks fit the 160 KB LDS of a CU

__device__ __forceinline__ int kwave(int l, int n) { return l <= n / 2 ? l : l - n; }

__device__ __forceinline__ bool dealias_zero(const Grid &g, int lx, int ly, int lz)
{
    if (g.basis == 1) {
        // SFF_PENCIL::Dealias (SFF_pencil_basic.cc:94-97): lx in [2Nx/3+1, Nx-1],
        // ly in [Ny/3+1, 2Ny/3-1], lz in [Nz/3+1, Nz/2] (empty y/z ranges
        // clamped); the 2/3 mode zeroes lx > 2Nx/3, |ky| > Ny/3, lz > Nz/3.
        if (g.dealias == 1) {
            int ylo = g.ny / 3 + 1, yhi = 2 * g.ny / 3 - 1;
            if (yhi < ylo) yhi = ylo;
            int zlo = g.nz / 3 + 1, zhi = g.nz / 2;
            if (zhi < zlo) zhi = zlo;
            return lx >= 2 * g.nx / 3 + 1 && (ly >= ylo && ly <= yhi) && (lz >= zlo && lz <= zhi);
        }
        if (g.dealias == 2) return lx > 2 * g.nx / 3 || abs(kwave(ly, g.ny)) > g.ny / 3 || lz > g.nz / 3;
        return false;
    }
    if (g.dealias == 1) {
        // FFF_PENCIL::Dealias (FFF_pencil_basic.cc:87-90) via Assign_sub_array
        // (:228-259): intersection of lx in [Nx/3+1, 2Nx/3-1],
        // ly in [Ny/3+1, 2Ny/3-1], lz in [Nz/3+1, Nz/2]; an empty y/z range is
        // clamped to its first index.
        int ylo = g.ny / 3 + 1, yhi = 2 * g.ny / 3 - 1;
        if (yhi < ylo) yhi = ylo;
        int zlo = g.nz / 3 + 1, zhi = g.nz / 2;
        if (zhi < zlo) zhi = zlo;
        return (lx >= g.nx / 3 + 1 && lx <= 2 * g.nx / 3 - 1) && (ly >= ylo && ly <= yhi) && (lz >= zlo && lz <= zhi);
    }
    if (g.dealias == 2) {
        int kx = abs(kwave(lx, g.nx)), ky = abs(kwave(ly, g.ny));
        return kx > g.nx / 3 || ky > g.ny / 3 || lz > g.nz / 3;
    }
    return false;
}

// Viscous factor exp(exponent) of Array_exp_ksqr: exponent f K^2, or with
// hyperviscosity (f + fh K^(q-2)) K^2 (universal_basic.cc:265-291), K^(q-2) =
// (K^2)^((q-2)/2) with integer division (q = 4: K^2), by repeated
// multiplication; q = 0: none.  ks = K^2.
__device__ __forceinline__ double visc_exp(double f, double fh, int q, double ks)
{
    if (q == 0) return exp(f * ks);
    const int n = q == 4 ? 1 : (q - 2) / 2;
    double kp = 1.0;
    for (int i = 0; i < n; i++) kp *= ks;
    return exp((f + fh * kp) * ks);
}

template <typename T>
__host__ __device__ constexpr int col_block(int L, int elem_bytes)
{
    int tpl = L / 8;
    int cb = 16;
    while (cb > 1 && (cb * tpl > 1024 || cb * lpad_len(L) * elem_bytes > kLdsBudget)) cb /= 2;
    return cb;
}

template <typename A, typename B>
__device__ __forceinline__ cpx<A> cvt(cpx<B> v) { return mk<A>((A)v.x, (A)v.y); }

// ---------------------------------------------------------------------------
// x pass helpers: staging of an (Nx x CB) column tile through LDS.
// ---------------------------------------------------------------------------
// Thread t moves column t % CB, rows t / CB + TPL i (i < 8): consecutive
// lanes cover the CB columns of a row.  The eight loads are issued before the
// first LDS write (a runtime-bounded loop would wait for each load in turn);
// a column outside the field (cols[c] < 0) loads column 0 and stages zeros.
template <int L, int CB, typename S, typename G>
__device__ __forceinline__ void stage_in(cpx<S> *lds, const cpx<G> *__restrict__ f, const int *cols, int ncol_total)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    const int c = threadIdx.x % CB, r0 = threadIdx.x / CB, col = cols[c];
    const cpx<G> *p = f + (size_t)r0 * ncol_total + max(col, 0);
    const size_t step = (size_t)TPL * ncol_total;
    cpx<G> x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = mk<G>(p[i * step].x, p[i * step].y);
#pragma unroll
    for (int i = 0; i < 8; i++) lds[c * LP + lpad(r0 + TPL * i)] = col >= 0 ? cvt<S>(x[i]) : mk<S>(0, 0);
}

// stage_in in two halves, so that several tiles' loads can be in flight
// together: stage_load fills x[8] (this thread's share), stage_put writes it
// to LDS (zeros for a column outside the field).
template <int L, int CB, typename G>
__device__ __forceinline__ void stage_load(cpx<G> *x, const cpx<G> *__restrict__ f, const int *cols, int ncol_total)
{
    constexpr int TPL = L / 8;
    const int c = threadIdx.x % CB, r0 = threadIdx.x / CB;
    const cpx<G> *p = f + (size_t)r0 * ncol_total + max(cols[c], 0);
    const size_t step = (size_t)TPL * ncol_total;
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = mk<G>(p[i * step].x, p[i * step].y);
}

template <int L, int CB, typename S, typename G>
__device__ __forceinline__ void stage_put(cpx<S> *lds, const cpx<G> *x, const int *cols)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    const int c = threadIdx.x % CB, r0 = threadIdx.x / CB;
    const bool in = cols[c] >= 0;
#pragma unroll
    for (int i = 0; i < 8; i++) lds[c * LP + lpad(r0 + TPL * i)] = in ? cvt<S>(x[i]) : mk<S>(0, 0);
}

template <int L, int CB, typename S, typename G>
__device__ __forceinline__ void stage_out(const cpx<S> *lds, cpx<G> *__restrict__ f, const int *cols, int ncol_total)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    const int c = threadIdx.x % CB, r0 = threadIdx.x / CB, col = cols[c];
    if (col < 0) return;
    cpx<G> *p = f + (size_t)r0 * ncol_total + col;
    const size_t step = (size_t)TPL * ncol_total;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const cpx<S> v = lds[c * LP + lpad(r0 + TPL * i)];
        p[i * step] = mk<G>((G)v.x, (G)v.y);
    }
}

// Tiles of the x pass.  Work fields store each ky row of kz with the padded
// stride nzp (a multiple of 128 bytes); the FP64 state in the boundary layout
// uses nzc = Nz/2+1.  A tile is CB consecutive columns of the padded (ly,lz)
// space (pad columns lz >= nzc are skipped), or -- patch tiles -- the
// kz = 0 and kz = Nz/2 columns of every ky.  colw = ly*nzp + lz indexes the
// work fields, cols = ly*nzc + lz the state.
//
// Blocks are dispatched to the 8 XCDs round-robin; xcd_tile() hands each XCD
// a contiguous range of tiles so that neighbouring tiles (which share the
// 128-byte lines of the unpadded state rows) run on the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int per = nb / 8;
    return (b % 8) * per + b / 8;
}

__host__ __device__ inline int x_tiles(const Grid &g, int CB, bool patch)
{
    return patch ? (2 * g.ny + CB - 1) / CB : (g.ny * g.nzp + CB - 1) / CB;
}

__host__ inline int x_grid(int ntiles) { return 8 * ((ntiles + 7) / 8); }

template <int CB>
__device__ __forceinline__ void tile_columns(int *colw, int *cols, const Grid &g, bool patch, int tile)
{
    if (threadIdx.x < CB) {
        const int c = threadIdx.x;
        int w = -1, st = -1;
        if (!patch) {
            const int cw = tile * CB + c;
            const int ly = cw / g.nzp, lz = cw % g.nzp;
            if (cw < g.ny * g.nzp && lz < g.nzc) { w = cw; st = ly * g.nzc + lz; }
        } else {
            const int p = tile * CB + c;
            if (p < 2 * g.ny) {
                const int ly = p >> 1, lz = (p & 1) ? g.nz / 2 : 0;
                w = ly * g.nzp + lz;
                st = ly * g.nzc + lz;
            }
        }
        colw[c] = w;
        cols[c] = st;
    }
}

// ---------------------------------------------------------------------------
// Direct-to-LDS staging of an x tile (global_load_lds_dwordx4): the L x CB
// tile of consecutive padded work columns cw0 .. cw0+CB-1 is moved in 1 KB
// chunks, one per wave instruction (64 lanes x 16 bytes, lane l landing at
// chunk + 16 l).  Within a chunk the layout is column-major, so the FFT's
// reads of one column are consecutive:
//   FP64 (16-byte elements): chunk = RPC = 64/CB rows x CB columns, lane
//        c*RPC + r%RPC holds element (r, c);
//   FP32 (8-byte elements): chunk = RPC2 = 128/CB rows x CB columns, lane
//        (c/2)*RPC2 + r%RPC2 holds the pair (r, c & ~1), (r, c | 1).
// Lanes of one instruction read CB adjacent columns of RPC (RPC2) rows: the
// same 64-byte row segments as a row-major staging.  The tile must fill whole
// waves and whole chunks (tile_dma_ok); the columns are always inside the
// padded rows, so pad columns are loaded too (and ignored).
// ---------------------------------------------------------------------------
template <typename T, int L, int CB>
__host__ __device__ constexpr bool tile_dma_ok()
{
    return (CB * (L / 8)) % 64 == 0 && (L * CB * (int)sizeof(cpx<T>)) % 1024 == 0 &&
           (sizeof(T) == 8 || CB % 2 == 0);
}

template <typename T, int L, int CB>
__device__ __forceinline__ void dma_tile(const cpx<T> *__restrict__ f, int cw0, int ncol_total, char *lds)
{
    constexpr int NWAVE = CB * (L / 8) / 64;
    constexpr int NCHUNK = L * CB * (int)sizeof(cpx<T>) / 1024;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = wave; k < NCHUNK; k += NWAVE) {
        const cpx<T> *src;
        if constexpr (sizeof(T) == 8) {
            constexpr int RPC = 64 / CB;
            src = f + (size_t)(k * RPC + lane % RPC) * ncol_total + cw0 + lane / RPC;
        } else {
            constexpr int RPC2 = 128 / CB;
            src = f + (size_t)(k * RPC2 + lane % RPC2) * ncol_total + cw0 + 2 * (lane / RPC2);
        }
        __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)(lds + k * 1024),
                                         16, 0, 0);
    }
}

// The same chunk layout for an FP64 tile of the state (boundary layout):
// column c of the tile is state column cols[c]; a column outside the state
// (cols[c] < 0, the padding of the work layout) loads state column 0 instead,
// whose values the caller ignores.
template <int L, int CB>
__device__ __forceinline__ void dma_state_tile(const cpx<double> *__restrict__ f, const int *cols, int ncol_total,
                                               char *lds)
{
    constexpr int NWAVE = CB * (L / 8) / 64;
    constexpr int NCHUNK = L * CB * 16 / 1024;
    constexpr int RPC = 64 / CB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = max(cols[lane / RPC], 0);
    for (int k = wave; k < NCHUNK; k += NWAVE) {
        const cpx<double> *src = f + (size_t)(k * RPC + lane % RPC) * ncol_total + col;
        __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)(lds + k * 1024),
                                         16, 0, 0);
    }
}

template <typename T, int L, int CB>
__device__ __forceinline__ cpx<T> dma_tile_read(const char *lds, int r, int c)
{
    if constexpr (sizeof(T) == 8) {
        constexpr int RPC = 64 / CB;
        return ((const cpx<T> *)lds)[(r / RPC) * 64 + c * RPC + (r % RPC)];
    } else {
        constexpr int RPC2 = 128 / CB;
        return *(const cpx<T> *)(lds + (r / RPC2) * 1024 + ((c / 2) * RPC2 + (r % RPC2)) * 16 + (c % 2) * 8);
    }
}

// Waits for this wave's direct-to-LDS loads (the barrier that follows makes
// every wave's loads visible): s_waitcnt vmcnt(0) with expcnt and lgkmcnt at
// their maxima.  The builtin (not inline asm) lets the compiler's wait
// insertion see that the loads have retired.
__device__ __forceinline__ void dma_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ int band_slot(const Band &b, int kx, int ky, int kz)
{
    if (kx < -b.kb || kx > b.kb || ky < -b.kb || ky > b.kb || kz > b.kb) return -1;
    return ((kx + b.kb) * (2 * b.kb + 1) + (ky + b.kb)) * (b.kb + 1) + kz;
}

__device__ __forceinline__ double kmag(const Grid &g, int kx, int ky, int kz)
{
    const double a = kx * g.kf[0], b = ky * g.kf[1], c = kz * g.kf[2];
    return sqrt(a * a + b * b + c * c);
}

// ---------------------------------------------------------------------------
// z-line transforms.  A real line x[0..N) (N = 2M) is packed as
// z[n] = x[2n] + i x[2n+1]; a complex line holds X[0..M].
// c2r: Z'[k] = (X[k] + conj X[M-k]) + i W^-k (X[k] - conj X[M-k]), with the
//      imaginary parts of X[0], X[M] dropped; z = IDFT_M(Z') (unscaled), so
//      x = sum_k X_herm[k] e^{+2 pi i k n / N}.
// r2c: Z = DFT_M(z); X[k] = ((Z[k] + conj Z[M-k]) - i W^k (Z[k] - conj Z[M-k]))/2.
// W = exp(-2 pi i / N); twn holds W^k for k = 0..M.
// ---------------------------------------------------------------------------
// The M+1 points of a complex line: v[s] = X[j + TPL*s], xm = X[M] (read by
// every thread of the line; member-wise copies keep v in registers).
template <typename T, int M>
__device__ __forceinline__ void zline_load(cpx<T> *v, cpx<T> &xm, const cpx<T> *__restrict__ X, int j)
{
    constexpr int TPL = M / 8;
#pragma unroll
    for (int s = 0; s < 8; s++) v[s] = mk<T>(X[j + TPL * s].x, X[j + TPL * s].y);
    xm = mk<T>(X[M].x, X[M].y);
}

// c2r of a line already loaded by zline_load (v in, samples out in v).
template <typename T, int M>
__device__ __forceinline__ void zline_c2r_from(cpx<T> *v, cpx<T> xm, cpx<T> *lds, const cpx<T> *__restrict__ twm,
                                               const cpx<T> *__restrict__ twn, int j)
{
    constexpr int TPL = M / 8;
    // stage the line (M+1 points) in LDS
#pragma unroll
    for (int s = 0; s < 8; s++) lds[lpad(j + TPL * s)] = mk<T>(v[s].x, v[s].y);
    if (j == 0) lds[lpad(M)] = xm;
    line_sync<M / 8>();
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int k = j + TPL * s;
        cpx<T> xk = lds[lpad(k)];
        cpx<T> xm = lds[lpad(M - k)];
        if (k == 0) { xk.y = 0; xm.y = 0; }
        const cpx<T> a = xk + conj(xm);
        const cpx<T> d = xk - conj(xm);
        const cpx<T> w = twn[k];  // W^k ; we need W^-k = conj(w)
        const cpx<T> e = mk<T>(d.x * w.x + d.y * w.y, d.y * w.x - d.x * w.y);
        v[s] = a + mk<T>(-e.y, e.x);
    }
    line_sync<M / 8>();
    fft_line<M, +1>(v, lds, twm, j);
}

template <typename T, int M>
__device__ __forceinline__ void zline_c2r(cpx<T> *v, const cpx<T> *__restrict__ X, cpx<T> *lds,
                                          const cpx<T> *__restrict__ twm, const cpx<T> *__restrict__ twn, int j)
{
    cpx<T> xm;
    zline_load<T, M>(v, xm, X, j);
    zline_c2r_from<T, M>(v, xm, lds, twm, twn, j);
}

template <typename T, int M>
__device__ __forceinline__ void zline_r2c_post(cpx<T> *v, cpx<T> *__restrict__ X, cpx<T> *lds,
                                               const cpx<T> *__restrict__ twm, const cpx<T> *__restrict__ twn, int j,
                                               bool store)
{
    constexpr int TPL = M / 8;
    fft_line<M, -1>(v, lds, twm, j);
#pragma unroll
    for (int s = 0; s < 8; s++) lds[lpad(j + TPL * s)] = v[s];
    line_sync<M / 8>();
    const T h = (T)0.5;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int k = j + TPL * s;
        const cpx<T> zk = v[s];
        const cpx<T> zm = lds[lpad(k == 0 ? 0 : M - k)];
        const cpx<T> a = zk + conj(zm);
        const cpx<T> d = zk - conj(zm);
        const cpx<T> w = twn[k];
        const cpx<T> e = d * w;  // W^k d
        if (store) X[k] = mk<T>(h * (a.x + e.y), h * (a.y - e.x));
    }
    if (store && j == 0) X[M] = mk<T>(v[0].x - v[0].y, (T)0);
    line_sync<M / 8>();
}

struct ZArgs {
    void *in[6];
    void *out[9];
    void *tape[3];  // optional: real-space velocity lines (Nx*Ny lines of Nz/2 packed points)
    const void *twm, *twn;
    Grid g;
    int ncomp;
};

template <int M>
__host__ __device__ constexpr int z_lines_per_block()
{
    return (M / 8) >= 256 ? 1 : 256 / (M / 8);
}

// ---------------------------------------------------------------------------
// Transform lengths with compiled kernels, and the size dispatch of the
// launchers (lengths are compile-time).  x and y lines take L, z lines the
// half length M = Nz/2.  Power-of-two lengths in both precisions; lengths
// with radix-3/5 factors in FP64.
//
// Each launcher fills its argument block and calls a per-length function
// (NAME_len<T, L>) that launches the kernels of that length.  The per-length
// functions are instantiated in length units -- the same source compiled
// once per unit with TG_LEN_UNIT = 1..4 (1: power-of-two FP64, 2:
// power-of-two FP32, 3 and 4: radix-3/5 lengths FP64) -- so that the kernel
// instances compile in parallel; unit 0 (the plain compilation) holds the
// launchers and declares the per-length functions extern.
// ---------------------------------------------------------------------------
#define TG_L_POW2(F, T, P) F(8, T, P) F(16, T, P) F(32, T, P) F(64, T, P) F(128, T, P) F(256, T, P) F(512, T, P) \
    F(1024, T, P)
#define TG_L_ODDA(F, T, P) F(24, T, P) F(40, T, P) F(48, T, P) F(80, T, P) F(96, T, P) F(120, T, P) F(160, T, P) \
    F(192, T, P)
#define TG_L_ODDB(F, T, P) F(240, T, P) F(320, T, P) F(384, T, P) F(480, T, P) F(640, T, P) F(768, T, P) F(960, T, P)
#define TG_M_POW2(F, T, P) F(8, T, P) F(16, T, P) F(32, T, P) F(64, T, P) F(128, T, P) F(256, T, P) F(512, T, P)
#define TG_M_ODDA(F, T, P) F(24, T, P) F(40, T, P) F(48, T, P) F(80, T, P) F(96, T, P) F(120, T, P) F(160, T, P) \
    F(192, T, P)
#define TG_M_ODDB(F, T, P) F(240, T, P) F(320, T, P) F(384, T, P) F(480, T, P)

#define TG_UNIT_1(FL, FM, P) TG_L_POW2(FL, double, P) TG_M_POW2(FM, double, P)
#define TG_UNIT_2(FL, FM, P) TG_L_POW2(FL, float, P) TG_M_POW2(FM, float, P)
#define TG_UNIT_3(FL, FM, P) TG_L_ODDA(FL, double, P) TG_M_ODDA(FM, double, P)
#define TG_UNIT_4(FL, FM, P) TG_L_ODDB(FL, double, P) TG_M_ODDB(FM, double, P)
#define TG_ALL_UNITS(FL, FM, P) TG_UNIT_1(FL, FM, P) TG_UNIT_2(FL, FM, P) TG_UNIT_3(FL, FM, P) TG_UNIT_4(FL, FM, P)
#define TG_CAT_(a, b) a##b
#define TG_CAT(a, b) TG_CAT_(a, b)
// the instantiations of this compilation's length unit
#define TG_THIS_UNIT(FL, FM) TG_CAT(TG_UNIT_, TG_LEN_UNIT)(FL, FM, )
#define TG_NO_LENGTHS(N, T, P)

#ifndef TG_LEN_UNIT
#define TG_LEN_UNIT 0
#endif

#define TG_CASE_P2(N, T, CALL) \
    case N: { constexpr int LL = N, MM = N; (void)LL; (void)MM; CALL; } break;
#define TG_CASE_ODD(N, T, CALL)                        \
    case N: {                                          \
        constexpr int LL = N, MM = N;                  \
        (void)LL;                                      \
        (void)MM;                                      \
        if constexpr (sizeof(T) == 8) { CALL; }        \
        else return hipErrorInvalidValue;              \
    } break;
// CALL (no top-level commas: parenthesise template argument lists) is
// compiled once per length with the constant LL (x/y lines) or MM (z half
// lines); a length without kernels returns hipErrorInvalidValue.
#define TG_DISPATCH_L(T, Lval, CALL)                                                                  \
    switch (Lval) {                                                                                   \
        TG_L_POW2(TG_CASE_P2, T, CALL) TG_L_ODDA(TG_CASE_ODD, T, CALL) TG_L_ODDB(TG_CASE_ODD, T, CALL) \
        default: return hipErrorInvalidValue;                                                         \
    }
#define TG_DISPATCH_M(T, Mval, CALL)                                                                  \
    switch (Mval) {                                                                                   \
        TG_M_POW2(TG_CASE_P2, T, CALL) TG_M_ODDA(TG_CASE_ODD, T, CALL) TG_M_ODDB(TG_CASE_ODD, T, CALL) \
        default: return hipErrorInvalidValue;                                                         \
    }

// Host checks of a length against the compiled ones (x/y lines L, z half
// lines M; FP32 takes powers of two only).
#define TG_OK_P2(N, T, P) case N: return true;
#define TG_OK_ODD(N, T, P) case N: return !fp32;
inline bool l_compiled(int n, bool fp32)
{
    switch (n) {
        TG_L_POW2(TG_OK_P2, , ) TG_L_ODDA(TG_OK_ODD, , ) TG_L_ODDB(TG_OK_ODD, , )
        default: return false;
    }
}
inline bool m_compiled(int n, bool fp32)
{
    switch (n) {
        TG_M_POW2(TG_OK_P2, , ) TG_M_ODDA(TG_OK_ODD, , ) TG_M_ODDB(TG_OK_ODD, , )
        default: return false;
    }
}
#undef TG_OK_P2
#undef TG_OK_ODD

}  // namespace tg
