// tg_mhd.hip -- MHD (Elsasser) stage kernels.
//
// The nonlinear term of MHD (compute_nlin_VF_MHD.cc:71-123) runs through the
// giver transform pipeline with u = Z- and w = Z+: the nine products
// Z-_j Z+_i give N+_i = sum_j d_j(Z-_j Z+_i) (giver form) and
// N-_j = sum_i d_i(Z+_i Z-_j) (transposed giver form).  The kernels here are
// the pointwise parts, one thread per mode of the FP64 boundary layout:
//   k_mhd_elsasser  Z+ = D(u) + D(b), Z- = D(u) - D(b)   (stage-1 input)
//   k_mhd_combine   N_u = (N+ + N-)/2, N_b = (N+ - N-)/2 (MHD.cc:150-163)
//                   for Compute_nlin
//   k_x_mhd         the fused x pass of an RK stage: x-forward of the eight
//                   trace-shifted products, N+ and N- together, their combination as
//                   above, RHS_u = -P N_u, RHS_b = -N_b (Tadvance_functions.cc:
//                   262-270), the Lawson RK4 stage for u (nu) and b (eta)
//                   (RK_VF_MHD.cc:126-190), and the x-inverse of the next
//                   stage's Z+- (one read of the products per stage; stage 4
//                   leaves the next step's stage-1 input)
#include <type_traits>

#include "tg_device.h"
#include "tg_kernels.h"
#include "tg_pass.h"
#include "tg_mhd.h"

namespace tg {

__device__ __forceinline__ int kwm(int l, int n) { return l <= n / 2 ? l : l - n; }

__device__ __forceinline__ bool dealias_zero_m(const Grid &g, int lx, int ly, int lz)
{
    if (g.dealias == 1) {
        int ylo = g.ny / 3 + 1, yhi = 2 * g.ny / 3 - 1;
        if (yhi < ylo) yhi = ylo;
        int zlo = g.nz / 3 + 1, zhi = g.nz / 2;
        if (zhi < zlo) zhi = zlo;
        return (lx >= g.nx / 3 + 1 && lx <= 2 * g.nx / 3 - 1) && (ly >= ylo && ly <= yhi) && (lz >= zlo && lz <= zhi);
    }
    if (g.dealias == 2) {
        int kx = abs(kwm(lx, g.nx)), ky = abs(kwm(ly, g.ny));
        return kx > g.nx / 3 || ky > g.ny / 3 || lz > g.nz / 3;
    }
    return false;
}

struct MhdArgs {
    double *V[3], *W[3];        // u, b (FP64 state)
    double *accV[3], *accW[3];  // RK accumulators
    const double *Np[3], *Nm[3];  // N+, N- (scaled, FP64)
    double *Zp[3], *Zm[3];      // next stage Elsasser inputs (FP64 spectral)
    Grid g;
    int stage;
    double dt, fh_u, f1_u, fh_b, f1_b;
    int *nonfinite;  // stage 4: set to 1 if the new state holds a non-finite value
};

__device__ __forceinline__ void lawson(int stage, const cpx<double> &v0in, const cpx<double> &r, double Eh, double E1,
                                       double dt, cpx<double> &ac, cpx<double> &vn)
{
    const double c_half = 0.5 * dt, c_one = 1.0 * dt;
    const double c_sixth = (1.0 / 6) * dt, c_third = (1.0 / 3) * dt;
    if (stage == 1) {
        ac = c_sixth * (E1 * r);
        vn = Eh * (v0in + c_half * r);
    } else if (stage == 2) {
        ac = ac + c_third * (Eh * r);
        vn = Eh * v0in + c_half * r;
    } else if (stage == 3) {
        ac = ac + c_third * (Eh * r);
        vn = Eh * (Eh * v0in + c_one * r);
    } else {
        ac = ac + c_sixth * r;
        vn = E1 * v0in + ac;
    }
}

#if TG_LEN_UNIT == 0
__global__ void k_mhd_elsasser(MhdArgs a)
{
    const Grid &g = a.g;
    const size_t n = (size_t)g.nx * g.ny * g.nzc;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int lz = i % g.nzc;
        const int ly = (i / g.nzc) % g.ny;
        const int lx = i / ((size_t)g.nzc * g.ny);
        const bool z = dealias_zero_m(g, lx, ly, lz);
        for (int c = 0; c < 3; c++) {
            cpx<double> u = ((cpx<double> *)a.V[c])[i], b = ((cpx<double> *)a.W[c])[i];
            if (z) u = b = mk<double>(0, 0);
            ((cpx<double> *)a.Zp[c])[i] = u + b;
            ((cpx<double> *)a.Zm[c])[i] = u - b;
        }
    }
}

// The same on the kz = 0 and kz = Nz/2 planes only (after the reality
// condition has rewritten them).
__global__ void k_mhd_elsasser_planes(MhdArgs a)
{
    const Grid &g = a.g;
    const int n = g.nx * g.ny;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 2 * n; t += gridDim.x * blockDim.x) {
        const int lz = t < n ? 0 : g.nz / 2, q = t < n ? t : t - n;
        const int lx = q / g.ny, ly = q % g.ny;
        const size_t i = ((size_t)lx * g.ny + ly) * g.nzc + lz;
        const bool z = dealias_zero_m(g, lx, ly, lz);
        for (int c = 0; c < 3; c++) {
            cpx<double> u = ((cpx<double> *)a.V[c])[i], b = ((cpx<double> *)a.W[c])[i];
            if (z) u = b = mk<double>(0, 0);
            ((cpx<double> *)a.Zp[c])[i] = u + b;
            ((cpx<double> *)a.Zm[c])[i] = u - b;
        }
    }
}

__global__ void k_mhd_combine(MhdArgs a)
{
    const Grid &g = a.g;
    const size_t n = (size_t)g.nx * g.ny * g.nzc;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; c++) {
            cpx<double> *pp = (cpx<double> *)a.V[c] + i, *mm = (cpx<double> *)a.W[c] + i;
            const cpx<double> p = *pp, m = *mm;
            *pp = 0.5 * (p + m);
            *mm = 0.5 * (p - m);
        }
}

#endif  // TG_LEN_UNIT == 0

// ---------------------------------------------------------------------------
// Fused MHD x pass (stages 1-4).  Tile staging as the NS x pass: CB columns
// of the padded (ly, lz) space by all Nx, the products arriving by
// direct-to-LDS loads into tile B while the previous one is transformed.
// ---------------------------------------------------------------------------
// giver-form product order of the eight trace-shifted products (P_xx - P_zz,
// P_yy - P_zz, the off-diagonal ones; P_zz itself is not formed): the stage
// uses N_u = (N+ + N-)/2 only projected, and the shift changes N+ and N- by
// the same gradient -i K P_zz, so N_b = (N+ - N-)/2 is unchanged
static __constant__ int kMhdOrder[8] = {0, 4, 3, 1, 7, 5, 2, 6};

template <typename T, int L, int CB>
__global__ void __launch_bounds__(CB *(L / 8), 1) k_x_mhd(XmhdArgs a)
{
    constexpr int TPL = L / 8;
    constexpr int LP = lpad_len(L);
    __shared__ __attribute__((aligned(16))) char rawA[CB * LP * sizeof(cpx<double>)];
    __shared__ __attribute__((aligned(16))) char rawB[CB * LP * sizeof(cpx<double>)];
    __shared__ int colw[CB], cols[CB];
    __shared__ cpx<T> twl[L];
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
    const double Ky = kwave(ly, g.ny) * g.kf[1], Kz = lz * g.kf[2];
    for (int i = threadIdx.x; i < L; i += CB * TPL) twl[i] = ((const cpx<T> *)a.tw)[i];
    const cpx<T> *tw = twl;
    const int ncw = g.ny * g.nzp, ncs = g.ny * g.nzc;
    const T sc = (T)a.scale;
    const T Kyt = (T)Ky, Kzt = (T)Kz;

    // --- x-forward of the products P[3i+j] = FT[Z-_j Z+_i]: N+_i += i K_j q
    // (giver form), N-_j += i K_i q (transposed), branch-free over the
    // components (a zero K adds a signed zero) ---
    cpx<T> Np[3][8], Nm[3][8];
#pragma unroll
    for (int s = 0; s < 8; s++) Np[0][s] = Np[1][s] = Np[2][s] = Nm[0][s] = Nm[1][s] = Nm[2][s] = mk<T>(0, 0);
    constexpr bool DMA = tile_dma_ok<T, L, CB>();
    const int cw0 = tile * CB;
    if constexpr (DMA) dma_tile<T, L, CB>((const cpx<T> *)a.P[kMhdOrder[0]], cw0, ncw, rawB);
#pragma unroll 1
    for (int pi = 0; pi < 8; pi++) {
        const int p = kMhdOrder[pi];
        const int pii = p / 3, pjj = p % 3;
        int jj = j;
        asm volatile("" : "+v"(jj));
        cpx<T> v[8];
        if constexpr (DMA) {
            dma_wait();
            lds_barrier();
#pragma unroll
            for (int s = 0; s < 8; s++) v[s] = dma_tile_read<T, L, CB>(rawB, jj + TPL * s, c);
            lds_barrier();
            if (pi + 1 < 8) dma_tile<T, L, CB>((const cpx<T> *)a.P[kMhdOrder[pi + 1]], cw0, ncw, rawB);
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
            const T Kj = pjj == 0 ? Kx : (pjj == 1 ? Kyt : Kzt);
            const T Ki = pii == 0 ? Kx : (pii == 1 ? Kyt : Kzt);
            Np[0][s] = Np[0][s] + mul_i(pii == 0 ? Kj : (T)0, q);
            Np[1][s] = Np[1][s] + mul_i(pii == 1 ? Kj : (T)0, q);
            Np[2][s] = Np[2][s] + mul_i(pii == 2 ? Kj : (T)0, q);
            Nm[0][s] = Nm[0][s] + mul_i(pjj == 0 ? Ki : (T)0, q);
            Nm[1][s] = Nm[1][s] + mul_i(pjj == 1 ? Ki : (T)0, q);
            Nm[2][s] = Nm[2][s] + mul_i(pjj == 2 ? Ki : (T)0, q);
        }
        if constexpr (!DMA) lds_barrier();
    }
    if constexpr (DMA) lds_barrier_async();

    // --- N_u = (N+ + N-)/2, N_b = (N+ - N-)/2, pressure on N_u, right-hand
    // sides RHS_u = -(N_u + i K p), RHS_b = -N_b (in place: Np <- RHS_u,
    // Nm <- RHS_b) ---
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int lx = j + TPL * s;
        const double K0 = kwave(lx, g.nx) * g.kf[0];
        cpx<double> nu[3], nb[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const cpx<double> pp = cvt<double>(Np[k][s]), mm = cvt<double>(Nm[k][s]);
            nu[k] = 0.5 * (pp + mm);
            nb[k] = 0.5 * (pp - mm);
        }
        const double kxs = K0 * K0;
        const double ks = (kxs + Ky * Ky) + Kz * Kz;
        cpx<double> div = mul_i(K0, nu[0]);
        div = div + mul_i(Ky, nu[1]);
        div = div + mul_i(Kz, nu[2]);
        cpx<double> pr = mk<double>(0, 0);
        if (!(lx == 0 && ly == 0 && lz == 0)) pr = mk<double>(div.x / ks, div.y / ks);
        const double Kc[3] = {K0, Ky, Kz};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const cpx<double> ru = nu[k] + mul_i(Kc[k], pr);
            Np[k][s] = cvt<T>(mk<double>(-ru.x, -ru.y));
            Nm[k][s] = cvt<T>(mk<double>(-nb[k].x, -nb[k].y));
        }
    }

    // --- Lawson stage for u (nu) and b (eta), next Z+- (stages 1-3) ---
    const int stage = a.stage;
    bool bad = false;
    auto comp = [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        cpx<double> un[8];
        // u
        stage_in<L, CB>(ldd, (const cpx<double> *)a.V[K], cols, ncs);
        if (stage >= 2) stage_in<L, CB>(ldb, (const cpx<double> *)a.accV[K], colw, ncw);
        lds_barrier();
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int lx = j + TPL * s;
            const int li = c * LP + lpad(lx);
            const double K0 = kwave(lx, g.nx) * g.kf[0];
            const double ks = (K0 * K0 + Ky * Ky) + Kz * Kz;
            cpx<double> v0 = ldd[li];
            if (stage == 1 && dealias_zero(g, lx, ly, lz)) v0 = mk<double>(0, 0);
            cpx<double> ac = stage >= 2 ? ldb[li] : mk<double>(0, 0), vn;
            lawson(stage, v0, cvt<double>(Np[K][s]), exp(a.fh_u * ks), exp(a.f1_u * ks), a.dt, ac, vn);
            ldb[li] = mk<double>(stage == 4 ? vn.x : ac.x, stage == 4 ? vn.y : ac.y);
            if (stage == 4 && !(isfinite(vn.x) && isfinite(vn.y))) bad = true;
            un[s] = vn;
        }
        lds_barrier();
        if (stage == 4) stage_out<L, CB>(ldb, (cpx<double> *)a.V[K], cols, ncs);
        else stage_out<L, CB>(ldb, (cpx<double> *)a.accV[K], colw, ncw);
        lds_barrier();
        // b
        stage_in<L, CB>(ldd, (const cpx<double> *)a.W[K], cols, ncs);
        if (stage >= 2) stage_in<L, CB>(ldb, (const cpx<double> *)a.accW[K], colw, ncw);
        lds_barrier();
        cpx<T> zp[8], zm[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int lx = j + TPL * s;
            const int li = c * LP + lpad(lx);
            const double K0 = kwave(lx, g.nx) * g.kf[0];
            const double ks = (K0 * K0 + Ky * Ky) + Kz * Kz;
            const bool z = dealias_zero(g, lx, ly, lz);
            cpx<double> w0 = ldd[li];
            if (stage == 1 && z) w0 = mk<double>(0, 0);
            cpx<double> ac = stage >= 2 ? ldb[li] : mk<double>(0, 0), bn;
            lawson(stage, w0, cvt<double>(Nm[K][s]), exp(a.fh_b * ks), exp(a.f1_b * ks), a.dt, ac, bn);
            ldb[li] = mk<double>(stage == 4 ? bn.x : ac.x, stage == 4 ? bn.y : ac.y);
            if (stage == 4 && !(isfinite(bn.x) && isfinite(bn.y))) bad = true;
            cpx<double> u = un[s];
            if (z) u = bn = mk<double>(0, 0);
            zp[s] = cvt<T>(u + bn);
            zm[s] = cvt<T>(u - bn);
        }
        lds_barrier();
        if (stage == 4) stage_out<L, CB>(ldb, (cpx<double> *)a.W[K], cols, ncs);
        else stage_out<L, CB>(ldb, (cpx<double> *)a.accW[K], colw, ncw);
        lds_barrier();
        // x-inverse of the next stage input (stage 4: of the next step's
        // first stage, D(u) +- D(b) of the new state): Z+ into U[3+K], Z- into U[K]
        fft_line<L, +1>(zp, ldt + c * LP, tw, j);
#pragma unroll
        for (int s = 0; s < 8; s++) ldt[c * LP + lpad(j + TPL * s)] = zp[s];
        lds_barrier();
        stage_out<L, CB>(ldt, (cpx<T> *)a.U[3 + K], colw, ncw);
        lds_barrier();
        fft_line<L, +1>(zm, ldt + c * LP, tw, j);
#pragma unroll
        for (int s = 0; s < 8; s++) ldt[c * LP + lpad(j + TPL * s)] = zm[s];
        lds_barrier();
        stage_out<L, CB>(ldt, (cpx<T> *)a.U[K], colw, ncw);
        lds_barrier();
    };
    comp(std::integral_constant<int, 0>{});
    comp(std::integral_constant<int, 1>{});
    comp(std::integral_constant<int, 2>{});
    if (bad && a.nonfinite) *a.nonfinite = 1;
}

template <typename T, int LL>
hipError_t x_mhd_len(const XmhdArgs &a, hipStream_t st)
{
    constexpr int CB = col_block<T>(LL, 2 * sizeof(cpx<double>));
    hipLaunchKernelGGL((k_x_mhd<T, LL, CB>), dim3(x_grid(x_tiles(a.g, CB, false))), dim3(CB * (LL / 8)), 0, st, a);
    return hipGetLastError();
}

#define TG_MHD_L(N, T, P) P template hipError_t x_mhd_len<T, N>(const XmhdArgs &, hipStream_t);

#if TG_LEN_UNIT == 0
TG_ALL_UNITS(TG_MHD_L, TG_NO_LENGTHS, extern)

template <typename T>
hipError_t launch_x_mhd(const XmhdArgs &a, hipStream_t st)
{
    TG_DISPATCH_L(T, a.g.nx, return (x_mhd_len<T, LL>)(a, st));
    return hipErrorInvalidValue;
}

template hipError_t launch_x_mhd<double>(const XmhdArgs &, hipStream_t);
template hipError_t launch_x_mhd<float>(const XmhdArgs &, hipStream_t);

static MhdArgs mhd_args(const Grid &g, double *const *V, double *const *W)
{
    MhdArgs a{};
    for (int c = 0; c < 3; c++) {
        a.V[c] = V[c];
        a.W[c] = W[c];
    }
    a.g = g;
    return a;
}

hipError_t launch_mhd_elsasser(const Grid &g, double *const *V, double *const *W, double *const *Zp,
                               double *const *Zm, hipStream_t st)
{
    MhdArgs a = mhd_args(g, V, W);
    for (int c = 0; c < 3; c++) {
        a.Zp[c] = Zp[c];
        a.Zm[c] = Zm[c];
    }
    hipLaunchKernelGGL(k_mhd_elsasser, dim3(4096), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_mhd_elsasser_planes(const Grid &g, double *const *V, double *const *W, double *const *Zp,
                                      double *const *Zm, hipStream_t st)
{
    MhdArgs a = mhd_args(g, V, W);
    for (int c = 0; c < 3; c++) {
        a.Zp[c] = Zp[c];
        a.Zm[c] = Zm[c];
    }
    const int n = 2 * g.nx * g.ny;
    hipLaunchKernelGGL(k_mhd_elsasser_planes, dim3((n + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_mhd_combine(const Grid &g, double *const *NU, double *const *NW, hipStream_t st)
{
    MhdArgs a = mhd_args(g, NU, NW);
    hipLaunchKernelGGL(k_mhd_combine, dim3(4096), dim3(256), 0, st, a);
    return hipGetLastError();
}

#else
TG_THIS_UNIT(TG_MHD_L, TG_NO_LENGTHS)
#endif  // TG_LEN_UNIT

}  // namespace tg
