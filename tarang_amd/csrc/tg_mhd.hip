// tg_mhd.hip -- MHD (Elsasser) stage kernels.
//
// The nonlinear term of MHD (compute_nlin_VF_MHD.cc:71-123) runs through the
// giver transform pipeline with u = Z- and w = Z+: the nine products
// Z-_j Z+_i give N+_i = sum_j d_j(Z-_j Z+_i) (giver form) and
// N-_j = sum_i d_i(Z+_i Z-_j) (transposed giver form).  The kernels here are
// the pointwise parts, one thread per mode of the FP64 boundary layout:
//   k_mhd_elsasser  Z+ = D(u) + D(b), Z- = D(u) - D(b)   (stage-1 input)
//   k_mhd_rk        N_u = (N+ + N-)/2, N_b = (N+ - N-)/2 (MHD.cc:150-163),
//                   RHS_u = -P N_u, RHS_b = -N_b (Tadvance_functions.cc:
//                   262-270), the Lawson RK4 stage for u (nu) and b (eta)
//                   (RK_VF_MHD.cc:126-190), and the next stage's Z+-.
#include "tg_device.h"
#include "tg_kernels.h"
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

__global__ void k_mhd_rk(MhdArgs a)
{
    const Grid &g = a.g;
    const size_t n = (size_t)g.nx * g.ny * g.nzc;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int lz = i % g.nzc;
        const int ly = (i / g.nzc) % g.ny;
        const int lx = i / ((size_t)g.nzc * g.ny);
        const double K[3] = {kwm(lx, g.nx) * g.kf[0], kwm(ly, g.ny) * g.kf[1], lz * g.kf[2]};
        const double kxs = K[0] * K[0];
        const double ks = (kxs + K[1] * K[1]) + K[2] * K[2];
        cpx<double> nu[3], nb[3];
        for (int c = 0; c < 3; c++) {
            const cpx<double> p = ((const cpx<double> *)a.Np[c])[i], m = ((const cpx<double> *)a.Nm[c])[i];
            nu[c] = 0.5 * (p + m);
            nb[c] = 0.5 * (p - m);
        }
        // pressure on N_u only (Pressure::Compute_pressure(U), Pressure.cc:56-72)
        cpx<double> div = mul_i(K[0], nu[0]);
        div = div + mul_i(K[1], nu[1]);
        div = div + mul_i(K[2], nu[2]);
        cpx<double> pr = mk<double>(0, 0);
        if (!(lx == 0 && ly == 0 && lz == 0)) pr = mk<double>(div.x / ks, div.y / ks);
        const bool z = dealias_zero_m(g, lx, ly, lz);
        const double Ehu = exp(a.fh_u * ks), E1u = exp(a.f1_u * ks);
        const double Ehb = exp(a.fh_b * ks), E1b = exp(a.f1_b * ks);
        for (int c = 0; c < 3; c++) {
            cpx<double> ru = nu[c] + mul_i(K[c], pr);
            ru = mk<double>(-ru.x, -ru.y);
            const cpx<double> rb = mk<double>(-nb[c].x, -nb[c].y);
            cpx<double> u0 = ((cpx<double> *)a.V[c])[i], b0 = ((cpx<double> *)a.W[c])[i];
            if (a.stage == 1 && z) u0 = b0 = mk<double>(0, 0);  // state dealiased in place at stage 1
            cpx<double> au = mk<double>(0, 0), ab = mk<double>(0, 0);
            if (a.stage >= 2) {
                au = ((cpx<double> *)a.accV[c])[i];
                ab = ((cpx<double> *)a.accW[c])[i];
            }
            cpx<double> un, bn;
            lawson(a.stage, u0, ru, Ehu, E1u, a.dt, au, un);
            lawson(a.stage, b0, rb, Ehb, E1b, a.dt, ab, bn);
            if (a.stage == 4) {
                ((cpx<double> *)a.V[c])[i] = un;
                ((cpx<double> *)a.W[c])[i] = bn;
                if (!(isfinite(un.x) && isfinite(un.y) && isfinite(bn.x) && isfinite(bn.y)) && a.nonfinite)
                    *a.nonfinite = 1;
            } else {
                ((cpx<double> *)a.accV[c])[i] = au;
                ((cpx<double> *)a.accW[c])[i] = ab;
            }
            if (z) un = bn = mk<double>(0, 0);
            ((cpx<double> *)a.Zp[c])[i] = un + bn;
            ((cpx<double> *)a.Zm[c])[i] = un - bn;
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

hipError_t launch_mhd_combine(const Grid &g, double *const *NU, double *const *NW, hipStream_t st)
{
    MhdArgs a = mhd_args(g, NU, NW);
    hipLaunchKernelGGL(k_mhd_combine, dim3(4096), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_mhd_rk(const Grid &g, int stage, double dt, double nu, double eta, double *const *V,
                         double *const *W, double *const *accV, double *const *accW, const double *const *Np,
                         const double *const *Nm, double *const *Zp, double *const *Zm, hipStream_t st,
                         int *nonfinite)
{
    MhdArgs a = mhd_args(g, V, W);
    for (int c = 0; c < 3; c++) {
        a.accV[c] = accV[c];
        a.accW[c] = accW[c];
        a.Np[c] = Np[c];
        a.Nm[c] = Nm[c];
        a.Zp[c] = Zp[c];
        a.Zm[c] = Zm[c];
    }
    a.stage = stage;
    a.dt = dt;
    a.fh_u = -nu * 0.5 * dt;
    a.f1_u = -nu * 1.0 * dt;
    a.fh_b = -eta * 0.5 * dt;
    a.f1_b = -eta * 1.0 * dt;
    a.nonfinite = nonfinite;
    hipLaunchKernelGGL(k_mhd_rk, dim3(8192), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace tg
