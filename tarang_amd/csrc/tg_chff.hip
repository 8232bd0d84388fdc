// tg_chff.hip -- the Chebyshev (ChFF) path: Chebyshev along x, Fourier along
// y and z.  Line kernels over the (ky, kz) columns of the spectral layout
// (lx, ly, lz), lz fastest: one thread per column (or per column, part and
// parity for the Helmholtz solves), so that neighbouring threads read
// neighbouring lz and every step of the sequential recurrences along x is a
// coalesced access.  All arithmetic FP64.
#include "tg_device.h"
#include "tg_kernels.h"
#include "tg_chff.h"

namespace tg {

namespace {

__device__ __forceinline__ int kw(int l, int n) { return l <= n / 2 ? l : l - n; }

constexpr double kChEps = 1e-15;  // myconstant.MYEPS (Global.cc:62)

// a_k <- w a_k for lx > 0 (Chebyshev weights over the cosine transform)
__global__ void k_cheb_scale(const double *src, double *dst, size_t plane2, size_t total2, double w)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total2; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = i < plane2 ? src[i] : w * src[i];
}

// Chebyshev derivative along x of E-double elements (E = 2 complex, 1 real):
// c_k b_k = b_{k+2} + 2 (k+1) a_{k+1}, k = N-2 .. 0, times kf.
template <int E>
__global__ void k_cheb_deriv(const double *A, double *B, int nx, size_t ncol, double kf)
{
    const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    double b1[E], b2[E];
    for (int e = 0; e < E; e++) b1[e] = b2[e] = 0.0;
    for (int e = 0; e < E; e++) B[((size_t)(nx - 1) * ncol + col) * E + e] = 0.0;
    for (int k = nx - 2; k >= 0; k--) {
        const size_t ia = ((size_t)(k + 1) * ncol + col) * E, ib = ((size_t)k * ncol + col) * E;
        for (int e = 0; e < E; e++) {
            const double bk = b2[e] + (2.0 * (k + 1)) * A[ia + e];
            b2[e] = b1[e];
            b1[e] = bk;
            B[ib + e] = (k == 0 ? 0.5 * bk : bk) * kf;
        }
    }
}

__device__ __forceinline__ int c_k(int k) { return k == 0 ? 2 : 1; }

// One parity of Helmholtz_real (helmholtz.cc:18-42) through
// Tridiagonal_solver (:155-231), Thual's algorithm: rows i = 1..m-1 of
// k = 2i + odd, first row all ones with value d0.  f(i) = R[k = 2i + odd],
// element stride E doubles, part offset `part`; X, Y: this thread's
// scratch, entry i at X[i * nthr].
__device__ void helm_parity(const double *R, double *U, int N, size_t ncol, size_t col, int E, int part, int odd,
                            double lambda, double d0, double *X, double *Y, size_t nthr)
{
    const int m = N / 2, n = m - 1;
    auto F = [&](int i) -> double {
        return i < m ? R[((size_t)(2 * i + odd) * ncol + col) * E + part] : 0.0;
    };
    auto rows = [&](int i, double &a, double &b, double &c, double &d) {
        const int k = 2 * i + odd;
        const int bk = k >= N - 2 ? 0 : 1, bk2 = k + 2 >= N - 2 ? 0 : 1;
        a = c_k(k - 2) * lambda / (4 * k * (k - 1));
        b = -(1 + lambda * bk / (2 * (k * k - 1)));
        c = lambda * bk2 / (4 * k * (k + 1));
        d = c_k(k - 2) * F(i - 1) / (4 * k * (k - 1)) - bk * F(i) / (2 * (k * k - 1)) + bk2 * F(i + 1) / (4 * k * (k + 1));
    };
    auto out = [&](int i, double v) { U[((size_t)(2 * i + odd) * ncol + col) * E + part] = v; };
    if (n <= 0) {
        out(0, d0);
        return;
    }
    double a, b, c, d;
    rows(n, a, b, c, d);
    double Xi = -a / b, Yi = d / b;  // X(n-1), Y(n-1)
    X[(size_t)(n - 1) * nthr] = Xi;
    Y[(size_t)(n - 1) * nthr] = Yi;
    for (int i = n - 1; i > 0; i--) {
        rows(i, a, b, c, d);
        const double den = b + c * Xi;
        const double Xn = -a / den;
        const double Yn = (d - (c * Yi)) / (b + c * Xi);
        Xi = Xn;
        Yi = Yn;
        X[(size_t)(i - 1) * nthr] = Xi;
        Y[(size_t)(i - 1) * nthr] = Yi;
    }
    double theta = 1, lamda = 0, THETA = 1, LAMDA = 0;
    for (int i = 1; i <= n; i++) {
        const double x = X[(size_t)(i - 1) * nthr], y = Y[(size_t)(i - 1) * nthr];
        theta = x * theta;
        lamda = (x * lamda) + y;
        THETA += theta;
        LAMDA += lamda;
    }
    double yv = (d0 - LAMDA) / THETA;
    out(0, yv);
    for (int i = 0; i < n; i++) {
        yv = X[(size_t)i * nthr] * yv + Y[(size_t)i * nthr];
        out(i + 1, yv);
    }
}

// Helmholtz_complex_full_array / Helmholtz_real_full_array (helmholtz.cc:
// 66-118): thread = (part, parity, column); lambda = Kperp^2 + lam_sup.
__global__ void k_helmholtz(HelmArgs a)
{
    const size_t ncol = (size_t)a.g.ny * a.g.nzc;
    const size_t nthr = ncol * 2 * a.E;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nthr) return;
    const size_t col = t % ncol;
    const int odd = (int)((t / ncol) % 2), part = (int)(t / (ncol * 2));
    const int ly = (int)(col / a.g.nzc), lz = (int)(col % a.g.nzc);
    const double Ky = kw(ly, a.g.ny) * a.g.kf[1], Kz = lz * a.g.kf[2];
    const double kp2 = Ky * Ky + Kz * Kz;
    const double vp = a.vp[part], vm = a.vm[part];
    const double d0 = odd ? (vp - vm) / 2 : (vp + vm) / 2;
    const int N = a.g.nx;
    if (a.pressure_switch && kp2 < kChEps) {
        // del^2 p = 0 at Kperp = 0: the linear profile through the wall values
        for (int i = 0; i < N / 2; i++)
            a.U[((size_t)(2 * i + odd) * ncol + col) * a.E + part] = i == 0 ? d0 : 0.0;
        return;
    }
    helm_parity(a.R, a.U, N, ncol, col, a.E, part, odd, kp2 + a.lam_sup, d0, a.X + t, a.Y + t, nthr);
}

// influence matrix (influence_matrix.cc:111-122)
__global__ void k_influence(const double *vxp, const double *vxm, double *M, int nx, size_t ncol, double kf)
{
    const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    double m00 = 0, m01 = 0, m10 = 0, m11 = 0;
    for (int kx = 0; kx <= nx - 2; kx += 2) {
        const double f1 = (double)kx * kx, f2 = (double)(kx + 1) * (kx + 1);
        const size_t i0 = (size_t)kx * ncol + col, i1 = (size_t)(kx + 1) * ncol + col;
        m00 += f1 * vxp[i0] + f2 * vxp[i1];
        m01 += f1 * vxm[i0] + f2 * vxm[i1];
        m10 += (-f1) * vxp[i0] + f2 * vxp[i1];
        m11 += (-f1) * vxm[i0] + f2 * vxm[i1];
    }
    M[4 * col + 0] = m00 * kf;
    M[4 * col + 1] = m01 * kf;
    M[4 * col + 2] = m10 * kf;
    M[4 * col + 3] = m11 * kf;
}

// Compute_pressure_Ux_chebyshev (influence_matrix.cc:137-180)
__global__ void k_influence_correct(InflArgs a)
{
    const size_t ncol = (size_t)a.ny * a.nzc;
    const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const double vpt = a.M[4 * col], vmt = a.M[4 * col + 1], vpb = a.M[4 * col + 2], vmb = a.M[4 * col + 3];
    const double det = vpt * vmb - vmt * vpb;
    cpx<double> dp = mk<double>(0, 0), dm = mk<double>(0, 0);
    const cpx<double> *V1 = (const cpx<double> *)a.V1;
    for (int kx = 0; kx <= a.nx - 2; kx += 2) {
        const double f1 = (double)kx * kx, f2 = (double)(kx + 1) * (kx + 1);
        const cpx<double> v0 = V1[(size_t)kx * ncol + col], v1 = V1[(size_t)(kx + 1) * ncol + col];
        dp = dp + (f1 * v0 + f2 * v1);
        dm = dm + ((-f1) * v0 + f2 * v1);
    }
    const cpx<double> vpp = a.kf * dp, vpm = a.kf * dm;
    cpx<double> delta_p = mk<double>(0, 0), delta_m = mk<double>(0, 0);
    if (fabs(det) > kChEps) {
        const cpx<double> np = vmt * vpm - vmb * vpp, nm = vpb * vpp - vpt * vpm;
        delta_p = mk<double>(np.x / det, np.y / det);
        delta_m = mk<double>(nm.x / det, nm.y / det);
    }
    const bool zero = col == 0;  // (ky, kz) = (0, 0): P and V1 set to zero
    cpx<double> *P = (cpx<double> *)a.P, *W = (cpx<double> *)a.V1;
    for (int k = 0; k < a.nx; k++) {
        const size_t i = (size_t)k * ncol + col;
        if (zero) {
            P[i] = mk<double>(0, 0);
            W[i] = mk<double>(0, 0);
            continue;
        }
        P[i] = P[i] + (a.pp[i] * delta_p + a.pm[i] * delta_m);
        W[i] = W[i] + (a.vxp[i] * delta_p + a.vxm[i] * delta_m);
    }
}

__global__ void k_real_product(const double *a, const double *b, double *c, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c[i] = a[i] * b[i];
}

// dst (+)= D (dir 0: the Chebyshev x-derivative held in D) or i K_dir P
__global__ void k_add_deriv(double *dst, const double *P, const double *D, Grid g, int dir, int assign)
{
    const size_t n = (size_t)g.nx * g.ny * g.nzc;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int lz = (int)(i % g.nzc), ly = (int)((i / g.nzc) % g.ny);
        cpx<double> t;
        if (dir == 0) {
            t = ((const cpx<double> *)D)[i];
        } else {
            const double K = dir == 1 ? kw(ly, g.ny) * g.kf[1] : lz * g.kf[2];
            t = mul_i(K, ((const cpx<double> *)P)[i]);
        }
        cpx<double> *o = (cpx<double> *)dst + i;
        *o = assign ? t : *o + t;
    }
}

// S <- ((S - V/dt) + grad_c p) / (-nu) (BDF1.cc:80-110); grad_x p from D
__global__ void k_bdf1_rhs(double *S, const double *V, const double *P, const double *D, Grid g, int comp, double dt,
                           double nu)
{
    const size_t n = (size_t)g.nx * g.ny * g.nzc;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int lz = (int)(i % g.nzc), ly = (int)((i / g.nzc) % g.ny);
        cpx<double> s = ((cpx<double> *)S)[i];
        const cpx<double> v = ((const cpx<double> *)V)[i];
        s = mk<double>(s.x - v.x / dt, s.y - v.y / dt);
        cpx<double> gp;
        if (comp == 0) gp = ((const cpx<double> *)D)[i];
        else gp = mul_i(comp == 1 ? kw(ly, g.ny) * g.kf[1] : lz * g.kf[2], ((const cpx<double> *)P)[i]);
        s = s + gp;
        ((cpx<double> *)S)[i] = mk<double>(s.x / (-nu), s.y / (-nu));
    }
}

__global__ void k_fill_double(double *a, size_t n, double v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_scale_real(double *a, size_t n, double s)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = a[i] / s;
}

dim3 grid_for(size_t n, int threads = 256)
{
    size_t b = (n + threads - 1) / threads;
    if (b > 8192) b = 8192;
    return dim3((unsigned)(b ? b : 1));
}

}  // namespace

hipError_t launch_cheb_scale(const Grid &g, const double *src, double *dst, double w, hipStream_t st)
{
    const size_t plane2 = (size_t)2 * g.ny * g.nzc, total2 = plane2 * g.nx;
    hipLaunchKernelGGL(k_cheb_scale, grid_for(total2), dim3(256), 0, st, src, dst, plane2, total2, w);
    return hipGetLastError();
}

hipError_t launch_cheb_deriv(const Grid &g, const double *A, double *B, bool complex_, hipStream_t st)
{
    const size_t ncol = (size_t)g.ny * g.nzc;
    dim3 grid((unsigned)((ncol + 127) / 128));
    if (complex_) hipLaunchKernelGGL(k_cheb_deriv<2>, grid, dim3(128), 0, st, A, B, g.nx, ncol, g.kf[0]);
    else hipLaunchKernelGGL(k_cheb_deriv<1>, grid, dim3(128), 0, st, A, B, g.nx, ncol, g.kf[0]);
    return hipGetLastError();
}

hipError_t launch_helmholtz(const HelmArgs &a, hipStream_t st)
{
    const size_t nthr = (size_t)a.g.ny * a.g.nzc * 2 * a.E;
    hipLaunchKernelGGL(k_helmholtz, dim3((unsigned)((nthr + 127) / 128)), dim3(128), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_influence(const Grid &g, const double *vxp, const double *vxm, double *M, hipStream_t st)
{
    const size_t ncol = (size_t)g.ny * g.nzc;
    hipLaunchKernelGGL(k_influence, dim3((unsigned)((ncol + 127) / 128)), dim3(128), 0, st, vxp, vxm, M, g.nx, ncol,
                       g.kf[0]);
    return hipGetLastError();
}

hipError_t launch_influence_correct(const InflArgs &a, hipStream_t st)
{
    const size_t ncol = (size_t)a.ny * a.nzc;
    hipLaunchKernelGGL(k_influence_correct, dim3((unsigned)((ncol + 127) / 128)), dim3(128), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_real_product(const double *a, const double *b, double *c, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_real_product, grid_for(n), dim3(256), 0, st, a, b, c, n);
    return hipGetLastError();
}

hipError_t launch_add_deriv(const Grid &g, double *dst, const double *P, const double *D, int dir, bool assign,
                            hipStream_t st)
{
    const size_t n = (size_t)g.nx * g.ny * g.nzc;
    hipLaunchKernelGGL(k_add_deriv, grid_for(n), dim3(256), 0, st, dst, P, D, g, dir, assign ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_bdf1_rhs(const Grid &g, double *S, const double *V, const double *P, const double *D, int comp,
                           double dt, double nu, hipStream_t st)
{
    const size_t n = (size_t)g.nx * g.ny * g.nzc;
    hipLaunchKernelGGL(k_bdf1_rhs, grid_for(n), dim3(256), 0, st, S, V, P, D, g, comp, dt, nu);
    return hipGetLastError();
}

hipError_t launch_fill_double(double *a, size_t n, double v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_double, grid_for(n), dim3(256), 0, st, a, n, v);
    return hipGetLastError();
}

hipError_t launch_scale_real(double *a, size_t n, double divisor, hipStream_t st)
{
    hipLaunchKernelGGL(k_scale_real, grid_for(n), dim3(256), 0, st, a, n, divisor);
    return hipGetLastError();
}

}  // namespace tg
