/*
 * tg_oracle_chff.c -- FP64 CPU restatement of the Chebyshev (ChFF) path:
 * Chebyshev along x, Fourier along y and z; the Chebyshev x-derivative; the
 * quasi-tridiagonal Helmholtz solver; the influence-matrix pressure/U_x
 * correction; the nonlinear term; the BDF1 time step.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker), like tg_oracle.c.
 *
 * The reference has the Chebyshev time-advance solvers
 * (trunk/lib/fluid/incompressible/time_advance/chebyshev/) but no ChFF basis
 * class (its slot is commented out, trunk/src/solvers/main.cc:86-87) and its
 * transform library (fftk) is absent, so the sample grid is this build's
 * choice: Chebyshev-Gauss points x_j = cos(pi (j + 1/2) / Nx), j = 0..Nx-1,
 * f(x_j) = sum_k a_k T_k(x_j), i.e. the SFF cosine transform of
 * tg_oracle_sff.c with a_0 = A_0, a_k = 2 A_k.  Arrays keep the FFF layout
 * (lx, ly, lz), lz fastest; the physical derivative is kfactor[0] d/dxi
 * (Compute_derivative_at_boundary multiplies by kfactor[1],
 * influence_matrix.cc:215-229).
 *
 * Pinning (tests/test_oracle_chff.py): transforms of Chebyshev polynomials
 * and of sampled functions against numpy's chebyshev module; the derivative
 * against numpy.polynomial.chebyshev.chebder; the Helmholtz solver against
 * a dense solve of the same collocation problem (u'' - lambda u = -f with
 * Dirichlet values) built with numpy; the BDF1 step: divergence-free
 * result with no-slip walls, the decay of a Stokes mode against the
 * linear solve.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tg_oracle_int.h"

typedef or_cplx cplx;

#define CH_MYEPS 1e-15

static inline size_t cix(const or_grid *g, int lx, int ly, int lz)
{
    return ((size_t)lx * g->ny + ly) * (size_t)(g->nz / 2 + 1) + lz;
}
static inline int cky(const or_grid *g, int ly) { return ly <= g->ny / 2 ? ly : ly - g->ny; }

/* ---- transforms ---- */
void or_chff_forward(const or_grid *g, const double *real_in, double *spec_out)
{
    or_sff_forward(g, OR_COS, real_in, spec_out);
    cplx *A = (cplx *)spec_out;
    const size_t plane = (size_t)g->ny * (g->nz / 2 + 1);
    for (size_t i = plane; i < (size_t)g->nx * plane; i++) A[i] *= 2.0;
}

void or_chff_inverse(const or_grid *g, const double *spec_in, double *real_out)
{
    const size_t plane = (size_t)g->ny * (g->nz / 2 + 1), tot = (size_t)g->nx * plane;
    cplx *A = (cplx *)malloc(sizeof(cplx) * tot);
    memcpy(A, spec_in, sizeof(cplx) * tot);
    for (size_t i = plane; i < tot; i++) A[i] *= 0.5;
    or_sff_inverse(g, OR_COS, (const double *)A, real_out);
    free(A);
}

/* d/dx along x of Chebyshev coefficients: c_k b_k = b_{k+2} + 2 (k+1) a_{k+1}
 * from k = Nx-2 down to 0 (b_{Nx-1} = b_{Nx} = 0), times kfactor[0]. */
static void cheb_deriv_line(const cplx *a, size_t stride, int n, double kf, cplx *b)
{
    cplx b2 = 0, b1 = 0;  /* b_{k+2}, b_{k+1} */
    b[(size_t)(n - 1) * stride] = 0;
    for (int k = n - 2; k >= 0; k--) {
        cplx bk = b2 + (2.0 * (k + 1)) * a[(size_t)(k + 1) * stride];
        b2 = b1;
        b1 = bk;
        b[(size_t)k * stride] = k == 0 ? 0.5 * bk : bk;
    }
    for (int k = 0; k < n; k++) b[(size_t)k * stride] *= kf;
}

void or_cheb_xderiv(const or_grid *g, const double *A_in, double *B_out)
{
    const cplx *A = (const cplx *)A_in;
    cplx *B = (cplx *)B_out;
    const size_t stride = (size_t)g->ny * (g->nz / 2 + 1);
    for (size_t col = 0; col < stride; col++) cheb_deriv_line(A + col, stride, g->nx, g->kfactor[0], B + col);
}

/* ---- Helmholtz solver (helmholtz.cc) ---- */
static inline int c_k(int k) { return k == 0 ? 2 : 1; }

/* Tridiagonal_solver (helmholtz.cc:155-231): the rows of one parity, first
 * row all ones (the boundary value d0), Thual's algorithm.  f and y have
 * N/2 entries; f(N/2) (read by the last row with a zero weight) is 0. */
static void tridiagonal(int N, const double *f, double *y, double lambda, double d0, int odd)
{
    const int m = N / 2, n = m - 1;
    double *a = (double *)calloc(m, sizeof(double)), *b = (double *)calloc(m, sizeof(double));
    double *c = (double *)calloc(m, sizeof(double)), *d = (double *)calloc(m, sizeof(double));
    b[0] = 1;
    c[0] = 1;
    d[0] = d0;
    for (int i = 1; i < m; i++) {
        const int k = 2 * i + odd;
        const int bk = k >= N - 2 ? 0 : 1, bk2 = k + 2 >= N - 2 ? 0 : 1;
        const double fi1 = i + 1 < m ? f[i + 1] : 0.0;
        a[i] = c_k(k - 2) * lambda / (4 * k * (k - 1));
        b[i] = -(1 + lambda * bk / (2 * (k * k - 1)));
        c[i] = lambda * bk2 / (4 * k * (k + 1));
        d[i] = c_k(k - 2) * f[i - 1] / (4 * k * (k - 1)) - bk * f[i] / (2 * (k * k - 1)) + bk2 * fi1 / (4 * k * (k + 1));
    }
    if (n <= 0) {
        y[0] = d[0];
        free(a); free(b); free(c); free(d);
        return;
    }
    double *X = (double *)malloc(sizeof(double) * n), *Y = (double *)malloc(sizeof(double) * n);
    X[n - 1] = -a[n] / b[n];
    Y[n - 1] = d[n] / b[n];
    for (int i = n - 1; i > 0; i--) {
        X[i - 1] = -a[i] / (b[i] + c[i] * X[i]);
        Y[i - 1] = (d[i] - (c[i] * Y[i])) / (b[i] + c[i] * X[i]);
    }
    double theta = 1, lamda = 0, THETA = 1, LAMDA = 0;
    for (int i = 1; i <= n; i++) {
        theta = X[i - 1] * theta;
        lamda = (X[i - 1] * lamda) + Y[i - 1];
        THETA += theta;
        LAMDA += lamda;
    }
    y[0] = (d[0] - LAMDA) / THETA;
    for (int i = 0; i < n; i++) y[i + 1] = X[i] * y[i] + Y[i];
    free(X); free(Y); free(a); free(b); free(c); free(d);
}

/* Helmholtz_real (helmholtz.cc:18-42): even and odd coefficients apart */
void or_helmholtz_line(int N, const double *f, double *x, double lambda, double vp, double vm)
{
    const int m = N / 2;
    double *h = (double *)malloc(sizeof(double) * m), *o = (double *)malloc(sizeof(double) * m);
    for (int i = 0; i < m; i++) h[i] = f[2 * i];
    tridiagonal(N, h, o, lambda, (vp + vm) / 2, 0);
    for (int i = 0; i < m; i++) x[2 * i] = o[i];
    for (int i = 0; i < m; i++) h[i] = f[2 * i + 1];
    tridiagonal(N, h, o, lambda, (vp - vm) / 2, 1);
    for (int i = 0; i < m; i++) x[2 * i + 1] = o[i];
    free(h); free(o);
}

/* Helmholtz_complex_full_array (helmholtz.cc:94-118), ChFF: per (ky, kz)
 * line, lambda = Kperp^2 + lambda_supplement, real and imaginary parts with
 * the real and imaginary boundary values. */
void or_helmholtz(const or_grid *g, const double *R_in, double *U_out, double lam_sup, const double vp[2],
                  const double vm[2])
{
    const int N = g->nx, nzc = g->nz / 2 + 1;
    const cplx *R = (const cplx *)R_in;
    cplx *U = (cplx *)U_out;
    double *f = (double *)malloc(sizeof(double) * N), *u = (double *)malloc(sizeof(double) * N);
    double *ui = (double *)malloc(sizeof(double) * N);
    for (int ly = 0; ly < g->ny; ly++)
        for (int lz = 0; lz < nzc; lz++) {
            const double Ky = cky(g, ly) * g->kfactor[1], Kz = lz * g->kfactor[2];
            const double lambda = (Ky * Ky + Kz * Kz) + lam_sup;
            for (int k = 0; k < N; k++) f[k] = creal(R[cix(g, k, ly, lz)]);
            or_helmholtz_line(N, f, u, lambda, vp[0], vm[0]);
            for (int k = 0; k < N; k++) f[k] = cimag(R[cix(g, k, ly, lz)]);
            or_helmholtz_line(N, f, ui, lambda, vp[1], vm[1]);
            for (int k = 0; k < N; k++) U[cix(g, k, ly, lz)] = u[k] + I * ui[k];
        }
    free(f); free(u); free(ui);
}

/* Helmholtz_real_full_array (helmholtz.cc:66-92) of a zero right-hand side
 * (the homogeneous solutions): real arrays in the (lx, ly, lz) layout. */
static void helmholtz_real_full(const or_grid *g, const double *R, double *U, int pressure_switch, double lam_sup,
                                double vp, double vm)
{
    const int N = g->nx, nzc = g->nz / 2 + 1;
    double *f = (double *)malloc(sizeof(double) * N), *u = (double *)malloc(sizeof(double) * N);
    for (int ly = 0; ly < g->ny; ly++)
        for (int lz = 0; lz < nzc; lz++) {
            const double Ky = cky(g, ly) * g->kfactor[1], Kz = lz * g->kfactor[2];
            const double kp2 = Ky * Ky + Kz * Kz;
            const size_t i0 = cix(g, 0, ly, lz), st = (size_t)g->ny * nzc;
            if (kp2 < CH_MYEPS && pressure_switch) {
                for (int k = 0; k < N; k++) U[i0 + k * st] = 0.0;
                U[i0] = (vp + vm) / 2;
                U[i0 + st] = (vp - vm) / 2;
                continue;
            }
            for (int k = 0; k < N; k++) f[k] = R[i0 + k * st];
            or_helmholtz_line(N, f, u, kp2 + lam_sup, vp, vm);
            for (int k = 0; k < N; k++) U[i0 + k * st] = u[k];
        }
    free(f); free(u);
}

static void cheb_deriv_real(const or_grid *g, const double *A, double *B)
{
    const int N = g->nx;
    const size_t st = (size_t)g->ny * (g->nz / 2 + 1);
    for (size_t col = 0; col < st; col++) {
        double b2 = 0, b1 = 0;
        B[col + (size_t)(N - 1) * st] = 0;
        for (int k = N - 2; k >= 0; k--) {
            const double bk = b2 + (2.0 * (k + 1)) * A[col + (size_t)(k + 1) * st];
            b2 = b1;
            b1 = bk;
            B[col + (size_t)k * st] = k == 0 ? 0.5 * bk : bk;
        }
        for (int k = 0; k < N; k++) B[col + (size_t)k * st] *= g->kfactor[0];
    }
}

/* Compute_homgeneous_soln_influence_matrix (influence_matrix.cc:74-133):
 * p+ (p(1) = 1, p(-1) = 0), p- (0, 1) with lambda = Kperp^2 (Kperp = 0:
 * the constant/linear solution), u_x+- from (D^2 - lambda') u = (dp/dx)/nu
 * with lambda' = Kperp^2 + 1/(dt nu) and no-slip values, and the 2x2
 * influence matrix of the wall derivatives of u_x+ and u_x-:
 * M[line] = {u+'(1), u-'(1), u+'(-1), u-'(-1)}.  Real fields (lx, ly, lz). */
void or_chff_influence(const or_grid *g, double nu, double dt, double *pp, double *pm, double *vxp, double *vxm,
                       double *M)
{
    const int N = g->nx, nzc = g->nz / 2 + 1;
    const size_t tot = (size_t)N * g->ny * nzc, st = (size_t)g->ny * nzc;
    double *z = (double *)calloc(tot, sizeof(double)), *t = (double *)malloc(sizeof(double) * tot);
    helmholtz_real_full(g, z, pp, 1, 0, 1.0, 0);
    cheb_deriv_real(g, pp, t);
    for (size_t i = 0; i < tot; i++) t[i] = t[i] / (-nu);
    helmholtz_real_full(g, t, vxp, 0, 1 / (dt * nu), 0, 0);
    helmholtz_real_full(g, z, pm, 1, 0, 0, 1.0);
    cheb_deriv_real(g, pm, t);
    for (size_t i = 0; i < tot; i++) t[i] = t[i] / (-nu);
    helmholtz_real_full(g, t, vxm, 0, 1 / (dt * nu), 0, 0);
    for (size_t col = 0; col < st; col++) {
        double m00 = 0, m01 = 0, m10 = 0, m11 = 0;
        for (int kx = 0; kx <= N - 2; kx += 2) {
            const double f1 = (double)kx * kx, f2 = (double)(kx + 1) * (kx + 1);
            m00 += f1 * vxp[col + kx * st] + f2 * vxp[col + (kx + 1) * st];
            m01 += f1 * vxm[col + kx * st] + f2 * vxm[col + (kx + 1) * st];
            m10 += (-f1) * vxp[col + kx * st] + f2 * vxp[col + (kx + 1) * st];
            m11 += (-f1) * vxm[col + kx * st] + f2 * vxm[col + (kx + 1) * st];
        }
        M[4 * col + 0] = m00 * g->kfactor[0];
        M[4 * col + 1] = m01 * g->kfactor[0];
        M[4 * col + 2] = m10 * g->kfactor[0];
        M[4 * col + 3] = m11 * g->kfactor[0];
    }
    free(z); free(t);
}

/* Compute_pressure_Ux_chebyshev (influence_matrix.cc:137-180) */
void or_chff_correct(const or_grid *g, const double *pp, const double *pm, const double *vxp, const double *vxm,
                     const double *M, double *P_io, double *V1_io)
{
    const int N = g->nx, nzc = g->nz / 2 + 1;
    const size_t st = (size_t)g->ny * nzc;
    cplx *P = (cplx *)P_io, *V1 = (cplx *)V1_io;
    for (size_t col = 0; col < st; col++) {
        const double vpt = M[4 * col], vmt = M[4 * col + 1], vpb = M[4 * col + 2], vmb = M[4 * col + 3];
        const double det = vpt * vmb - vmt * vpb;
        cplx dp = 0, dm = 0;
        for (int kx = 0; kx <= N - 2; kx += 2) {
            const double f1 = (double)kx * kx, f2 = (double)(kx + 1) * (kx + 1);
            dp += f1 * V1[col + kx * st] + f2 * V1[col + (kx + 1) * st];
            dm += -f1 * V1[col + kx * st] + f2 * V1[col + (kx + 1) * st];
        }
        const cplx vpp = g->kfactor[0] * dp, vpm = g->kfactor[0] * dm;
        cplx delta_p = 0, delta_m = 0;
        if (fabs(det) > CH_MYEPS) {
            delta_p = (vmt * vpm - vmb * vpp) / det;
            delta_m = (vpb * vpp - vpt * vpm) / det;
        }
        for (int k = 0; k < N; k++) {
            const size_t i = col + k * st;
            P[i] += delta_p * pp[i] + delta_m * pm[i];
            V1[i] += delta_p * vxp[i] + delta_m * vxm[i];
        }
    }
    for (int k = 0; k < N; k++) {
        P[cix(g, k, 0, 0)] = 0;
        V1[cix(g, k, 0, 0)] = 0;
    }
}

/* ---- nonlinear term: N_i = sum_j d_j (u_j u_i) in divergence form, x
 * derivatives by the Chebyshev recurrence (Compute_nlin_diag / _offdiag,
 * compute_diag_offdiag.cc:47-112, with the ChFF derivatives); no dealiasing */
void or_compute_nlin_chff(const or_grid *g, double *const V[3], double *N[3])
{
    const int nzc = g->nz / 2 + 1;
    const size_t nc = (size_t)g->nx * g->ny * nzc, nr = (size_t)g->nx * g->ny * (g->nz + 2);
    double *r[3], *pr = (double *)malloc(sizeof(double) * nr);
    cplx *P = (cplx *)malloc(sizeof(cplx) * nc), *D = (cplx *)malloc(sizeof(cplx) * nc);
    for (int c = 0; c < 3; c++) {
        r[c] = (double *)malloc(sizeof(double) * nr);
        or_chff_inverse(g, V[c], r[c]);
        memset(N[c], 0, sizeof(cplx) * nc);
    }
    /* products xx, yy, zz, xy, yz, zx; (i, j) pairs receiving d_j of the product */
    static const int pi_[6] = {0, 1, 2, 0, 1, 2}, pj_[6] = {0, 1, 2, 1, 2, 0};
    for (int p = 0; p < 6; p++) {
        const int a = pi_[p], b = pj_[p];
        for (size_t i = 0; i < nr; i++) pr[i] = r[a][i] * r[b][i];
        or_chff_forward(g, pr, (double *)P);
        /* N_a += d_b P; N_b += d_a P (once when a == b) */
        for (int t = 0; t < (a == b ? 1 : 2); t++) {
            const int dst = t == 0 ? a : b, dir = t == 0 ? b : a;
            cplx *Nd = (cplx *)N[dst];
            if (dir == 0) {
                or_cheb_xderiv(g, (const double *)P, (double *)D);
                for (size_t i = 0; i < nc; i++) Nd[i] += D[i];
            } else {
                for (int lx = 0; lx < g->nx; lx++)
                    for (int ly = 0; ly < g->ny; ly++)
                        for (int lz = 0; lz < nzc; lz++) {
                            const size_t i = cix(g, lx, ly, lz);
                            const double K = dir == 1 ? cky(g, ly) * g->kfactor[1] : lz * g->kfactor[2];
                            Nd[i] += I * K * P[i];
                        }
            }
        }
    }
    for (int c = 0; c < 3; c++) free(r[c]);
    free(pr); free(P); free(D);
}

/* BDF1 (BDF1.cc:63-121) without forcing: pressure from the divergence of N
 * (zero wall values), then each component from (D^2 - Kperp^2 - 1/(dt nu)) u
 * = (N_i - u_i/dt + d_i p)/(-nu) with no-slip walls, and the influence-matrix
 * correction of p and u_x (so that d u_x/dx = 0 at the walls).  P is the
 * pressure array (output). */
void or_bdf1_chff(const or_grid *g, double nu, double dt, double *V[3], double *P_out)
{
    const int N = g->nx, nzc = g->nz / 2 + 1;
    const size_t nc = (size_t)N * g->ny * nzc;
    double *NL[3], *pp, *pm, *vxp, *vxm, *M;
    for (int c = 0; c < 3; c++) NL[c] = (double *)malloc(sizeof(cplx) * nc);
    pp = (double *)malloc(sizeof(double) * nc);
    pm = (double *)malloc(sizeof(double) * nc);
    vxp = (double *)malloc(sizeof(double) * nc);
    vxm = (double *)malloc(sizeof(double) * nc);
    M = (double *)malloc(sizeof(double) * 4 * (size_t)g->ny * nzc);
    or_chff_influence(g, nu, dt, pp, pm, vxp, vxm, M);
    or_compute_nlin_chff(g, V, NL);
    cplx *div = (cplx *)malloc(sizeof(cplx) * nc), *D = (cplx *)malloc(sizeof(cplx) * nc);
    or_cheb_xderiv(g, NL[0], (double *)div);
    for (int lx = 0; lx < N; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const size_t i = cix(g, lx, ly, lz);
                div[i] += I * (cky(g, ly) * g->kfactor[1]) * ((cplx *)NL[1])[i] +
                          I * (lz * g->kfactor[2]) * ((cplx *)NL[2])[i];
            }
    const double zero[2] = {0, 0};
    or_helmholtz(g, (const double *)div, P_out, 0, zero, zero);
    const double lam = 1 / (dt * nu);
    cplx *Pc = (cplx *)P_out;
    for (int c = 0; c < 3; c++) {
        cplx *S = (cplx *)NL[c];
        const cplx *Vc = (const cplx *)V[c];
        for (size_t i = 0; i < nc; i++) S[i] -= Vc[i] / dt;
        if (c == 0) {
            or_cheb_xderiv(g, P_out, (double *)D);
            for (size_t i = 0; i < nc; i++) S[i] += D[i];
        } else {
            for (int lx = 0; lx < N; lx++)
                for (int ly = 0; ly < g->ny; ly++)
                    for (int lz = 0; lz < nzc; lz++) {
                        const size_t i = cix(g, lx, ly, lz);
                        const double K = c == 1 ? cky(g, ly) * g->kfactor[1] : lz * g->kfactor[2];
                        S[i] += I * K * Pc[i];
                    }
        }
        for (size_t i = 0; i < nc; i++) S[i] /= (-nu);
        or_helmholtz(g, NL[c], V[c], lam, zero, zero);
        if (c == 0) or_chff_correct(g, pp, pm, vxp, vxm, M, P_out, V[0]);
    }
    for (int c = 0; c < 3; c++) free(NL[c]);
    free(pp); free(pm); free(vxp); free(vxm); free(M); free(div); free(D);
}
