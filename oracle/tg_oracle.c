/*
 * tg_oracle.c -- FP64 CPU restatement of Tarang's pseudo-spectral hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Paths are relative to the reference checkout (anandogc/tarang).
 *
 * Pinning of this restatement (checks live in tests/test_oracle.py):
 *   or_fft1d, or_forward, or_inverse  -- numpy.fft golden fixtures
 *       (tests/golden/fft_*.npz), brute-force DFT, single-mode inputs,
 *       Parseval, round trip.
 *   or_dealias                         -- mode-count closed forms for the
 *       reference index ranges (FFF_pencil_basic.cc:87-90) and the 2/3 rule.
 *   or_strong_reality                  -- Hermitian plane after fill; the
 *       inverse of a filled field equals the inverse of its Hermitian part.
 *   or_compute_nlin                    -- triad closed form, Taylor-Green
 *       (u.grad)u at t=0, Beltrami P.N = 0, energy conservation (2/3 rule).
 *   or_compute_nlin_giver              -- linearity in the giver; giver = u
 *       reproduces or_compute_nlin.
 *   or_compute_rhs                     -- K.RHS = 0, closed-form projection.
 *   or_rk4_step / or_euler_step        -- N=0 gives exp(-nu K^2 t) exactly,
 *       4th-order convergence in dt, stage formulas on a single mode.
 *   or_energy, or_Sn                   -- Parseval against real space,
 *       single-mode values.
 *   or_shell_spectrum                  -- hand-computed 8^3 sums.
 *   or_flux, or_shell_to_shell         -- antisymmetry, zero sums,
 *       flux = sum of shell-to-shell blocks (2/3 rule, nu=0).
 *   or_init_field                      -- divergence free, Hermitian,
 *       exact total energy (this build's own seeded definition).
 * The reference itself cannot be built here (MPI, FFTW, Blitz++, yaml-cpp,
 * HDF5 and the external fftk transform library are all absent), so no
 * vector from a running reference pins any of these.
 */
#include "tg_oracle.h"
#include "tg_oracle_int.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* 1D mixed-radix FFT (recursive decimation in time, own implementation)     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n, nf;
    int f[64];
    cplx *w; /* w[k] = exp(-2 pi i k / n) */
} fftplan;

static int plan_init(fftplan *p, int n)
{
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
    for (int q = 3; q * q <= m; q += 2)
        while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
    if (m > 1) p->f[p->nf++] = m;
    p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    if (!p->w) return -1;
    for (int k = 0; k < n; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        p->w[k] = cos(a) + I * sin(a);
    }
    return 0;
}

static void plan_free(fftplan *p) { free(p->w); p->w = NULL; }

static void fft_rec(const fftplan *p, const cplx *in, int is, cplx *out, int n, int fi,
                    int wstride, int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    int r = p->f[fi];
    int m = n / r;
    for (int j = 0; j < r; j++)
        fft_rec(p, in + (size_t)j * is, is * r, out + (size_t)j * m, m, fi + 1, wstride * r, sign);
    const int N = p->n;
    /* twiddle index j*k*wstride < r*m*wstride = N: no reduction needed */
    if (r == 2) {
        for (int k = 0; k < m; k++) {
            cplx w = p->w[(size_t)k * wstride];
            if (sign > 0) w = conj(w);
            cplx a = out[k], b = out[k + m] * w;
            out[k] = a + b;
            out[k + m] = a - b;
        }
        return;
    }
    if (r == 4) {
        const cplx rot = sign > 0 ? I : -I; /* w_4^1 */
        for (int k = 0; k < m; k++) {
            cplx w1 = p->w[(size_t)k * wstride], w2 = p->w[(size_t)2 * k * wstride], w3 = p->w[(size_t)3 * k * wstride];
            if (sign > 0) { w1 = conj(w1); w2 = conj(w2); w3 = conj(w3); }
            cplx a0 = out[k], a1 = out[k + m] * w1, a2 = out[k + 2 * m] * w2, a3 = out[k + 3 * m] * w3;
            cplx b0 = a0 + a2, b1 = a0 - a2, b2 = a1 + a3, b3 = (a1 - a3) * rot;
            out[k] = b0 + b2;
            out[k + m] = b1 + b3;
            out[k + 2 * m] = b0 - b2;
            out[k + 3 * m] = b1 - b3;
        }
        return;
    }
    cplx t[64];
    cplx *tt = r <= 64 ? t : (cplx *)malloc(sizeof(cplx) * (size_t)r);
    for (int k = 0; k < m; k++) {
        for (int j = 0; j < r; j++) {
            cplx w = p->w[(size_t)j * k * wstride];
            if (sign > 0) w = conj(w);
            tt[j] = out[(size_t)j * m + k] * w;
        }
        for (int q = 0; q < r; q++) {
            cplx s = 0;
            for (int j = 0; j < r; j++) {
                long e = (long)((j * q) % r) * (N / r);
                cplx w = p->w[e];
                if (sign > 0) w = conj(w);
                s += tt[j] * w;
            }
            out[(size_t)q * m + k] = s;
        }
    }
    if (tt != t) free(tt);
}

/* x <- DFT(x) with exponent sign*2*pi*i*j*k/n; work has n entries. */
static void fft_exec(const fftplan *p, cplx *x, cplx *work, int sign)
{
    memcpy(work, x, sizeof(cplx) * (size_t)p->n);
    fft_rec(p, work, 1, x, p->n, 0, 1, sign);
}

int or_fft1d(double *data, int n, int sign)
{
    if (n <= 0) return -1;
    fftplan p;
    if (plan_init(&p, n)) return -1;
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    fft_exec(&p, (cplx *)data, work, sign);
    free(work);
    plan_free(&p);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3D FFF transforms (fftk "FFF": FFF_pencil_tr.cc:43-66)                    */
/* ------------------------------------------------------------------------ */

static inline size_t cidx(const or_grid *g, int lx, int ly, int lz)
{
    return ((size_t)lx * g->ny + ly) * (size_t)(g->nz / 2 + 1) + lz;
}

/* complex-to-complex transform along y of a (Nx,Ny,Nzc) array, unnormalised */
void or_int_fft_y(const or_grid *g, cplx *A, int sign)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    if (ny <= 1) return;
#pragma omp parallel
    {
        fftplan p;
        plan_init(&p, ny);
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)ny);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)ny);
#pragma omp for collapse(2) schedule(static)
        for (int lx = 0; lx < nx; lx++)
            for (int lz = 0; lz < nzc; lz++) {
                for (int ly = 0; ly < ny; ly++) line[ly] = A[cidx(g, lx, ly, lz)];
                fft_exec(&p, line, work, sign);
                for (int ly = 0; ly < ny; ly++) A[cidx(g, lx, ly, lz)] = line[ly];
            }
        free(line); free(work); plan_free(&p);
    }
}

/* complex-to-complex transform along x, unnormalised */
static void fft_x(const or_grid *g, cplx *A, int sign)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    if (nx <= 1) return;
#pragma omp parallel
    {
        fftplan p;
        plan_init(&p, nx);
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for collapse(2) schedule(static)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                for (int lx = 0; lx < nx; lx++) line[lx] = A[cidx(g, lx, ly, lz)];
                fft_exec(&p, line, work, sign);
                for (int lx = 0; lx < nx; lx++) A[cidx(g, lx, ly, lz)] = line[lx];
            }
        free(line); free(work); plan_free(&p);
    }
}

/* r2c along z of every (x, y) line of a real (Nx,Ny,Nz+2) array, unnormalised */
void or_int_z_r2c(const or_grid *g, const double *real_in, cplx *A)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nzc = nz / 2 + 1;
#pragma omp parallel
    {
        fftplan p;
        plan_init(&p, nz);
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
#pragma omp for collapse(2) schedule(static)
        for (int lx = 0; lx < nx; lx++)
            for (int ly = 0; ly < ny; ly++) {
                const double *r = real_in + ((size_t)lx * ny + ly) * (size_t)(nz + 2);
                for (int z = 0; z < nz; z++) line[z] = r[z];
                fft_exec(&p, line, work, -1);
                for (int kz = 0; kz < nzc; kz++) A[cidx(g, lx, ly, kz)] = line[kz];
            }
        free(line); free(work); plan_free(&p);
    }
}

/* c2r along z of every line (imaginary parts of kz = 0 and Nz/2 ignored), unnormalised */
void or_int_z_c2r(const or_grid *g, const cplx *A, double *real_out)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#pragma omp parallel
    {
        fftplan p;
        plan_init(&p, nz);
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
#pragma omp for collapse(2) schedule(static)
        for (int lx = 0; lx < nx; lx++)
            for (int ly = 0; ly < ny; ly++) {
                const cplx *c = A + cidx(g, lx, ly, 0);
                line[0] = creal(c[0]);
                for (int k = 1; k < nz / 2; k++) {
                    line[k] = c[k];
                    line[nz - k] = conj(c[k]);
                }
                line[nz / 2] = creal(c[nz / 2]);
                fft_exec(&p, line, work, +1);
                double *r = real_out + ((size_t)lx * ny + ly) * (size_t)(nz + 2);
                for (int z = 0; z < nz; z++) r[z] = creal(line[z]);
                r[nz] = 0.0;
                r[nz + 1] = 0.0;
            }
        free(line); free(work); plan_free(&p);
    }
}

/* Forward_transform (FFF_pencil_tr.cc:43-51): r2c along z, then y, x; 1/N^3. */
void or_forward(const or_grid *g, const double *real_in, double *spec_out)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nzc = nz / 2 + 1;
    cplx *A = (cplx *)spec_out;
    or_int_z_r2c(g, real_in, A);
    or_int_fft_y(g, A, -1);
    fft_x(g, A, -1);
    const double s = 1.0 / ((double)nx * (double)ny * (double)nz);
    const size_t tot = (size_t)nx * ny * nzc;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < tot; i++) A[i] *= s;
}

/* Inverse_transform (FFF_pencil_tr.cc:57-66): input preserved (the reference
 * copies to X_transform first, :59); x, y inverse, then c2r along z. */
void or_inverse(const or_grid *g, const double *spec_in, double *real_out)
{
    const size_t tot = (size_t)g->nx * g->ny * (g->nz / 2 + 1);
    cplx *A = (cplx *)malloc(sizeof(cplx) * tot);
    memcpy(A, spec_in, sizeof(cplx) * tot);
    fft_x(g, A, +1);
    or_int_fft_y(g, A, +1);
    or_int_z_c2r(g, A, real_out);
    free(A);
}

/* ------------------------------------------------------------------------ */
/* Basis helpers                                                            */
/* ------------------------------------------------------------------------ */

static inline int get_kx(const or_grid *g, int lx) { return lx <= g->nx / 2 ? lx : lx - g->nx; }
static inline int get_ky(const or_grid *g, int ly) { return ly <= g->ny / 2 ? ly : ly - g->ny; }

/* Kmagnitude, waveno_switch = true (FFF_pencil_inline.h:486-493) */
double or_kmag(const or_grid *g, int lx, int ly, int lz)
{
    double a = get_kx(g, lx) * g->kfactor[0];
    double b = get_ky(g, ly) * g->kfactor[1];
    double c = lz * g->kfactor[2];
    return sqrt(a * a + b * b + c * c);
}

static inline double ksqr(const or_grid *g, int lx, int ly, int lz)
{
    /* Array_divide_ksqr / Array_exp_ksqr order (universal_basic.cc:208-215) */
    double kx = get_kx(g, lx) * g->kfactor[0];
    double ky = get_ky(g, ly) * g->kfactor[1];
    double kz = lz * g->kfactor[2];
    double kxs = kx * kx;
    double kxys = kxs + ky * ky;
    return kxys + kz * kz;
}

/* Min_radius_outside (FFF_pencil_inline.h:497-504) */
int or_min_radius_outside(const or_grid *g)
{
    double a = (g->nx / 2) * g->kfactor[0], b = (g->ny / 2) * g->kfactor[1], c = (g->nz / 2) * g->kfactor[2];
    return (int)ceil(sqrt(a * a + b * b + c * c));
}

/* Max_radius_inside (FFF_pencil_inline.h:509-540), 3D branch */
int or_max_radius_inside(const or_grid *g)
{
    double k = fmin((g->nx / 2) * g->kfactor[0], (g->ny / 2) * g->kfactor[1]);
    k = fmin(k, (g->nz / 2) * g->kfactor[2]);
    return (int)k;
}

/* Multiplicity_factor (FFF_pencil_inline.h:583-602) */
double or_multiplicity(const or_grid *g, int lx, int ly, int lz)
{
    double f = lz > 0 ? 1.0 : 0.5;
    int kx = get_kx(g, lx), ky = get_ky(g, ly);
    if (kx == g->nx / 2 || (ky == g->ny / 2 && g->ny > 1)) return 2 * f;
    return f;
}

static int dealias_zero(const or_grid *g, int lx, int ly, int lz)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (g->dealias == OR_DEALIAS_REF_CORNER) {
        /* FFF_PENCIL::Dealias (FFF_pencil_basic.cc:87-90): intersection of the
         * index ranges lx in [Nx/3+1, 2Nx/3-1], ly in [Ny/3+1, 2Ny/3-1],
         * lz in [Nz/3+1, Nz/2]. Assign_sub_array (:228-259) clamps an empty
         * y/z range to its first index. */
        int ylo = ny / 3 + 1, yhi = 2 * ny / 3 - 1;
        if (yhi < ylo) yhi = ylo;
        int zlo = nz / 3 + 1, zhi = nz / 2;
        if (zhi < zlo) zhi = zlo;
        return (lx >= nx / 3 + 1 && lx <= 2 * nx / 3 - 1) && (ly >= ylo && ly <= yhi) &&
               (lz >= zlo && lz <= zhi);
    }
    if (g->dealias == OR_DEALIAS_TWOTHIRDS) {
        int kx = abs(get_kx(g, lx)), ky = abs(get_ky(g, ly));
        return kx > nx / 3 || ky > ny / 3 || lz > nz / 3;
    }
    return 0;
}

void or_dealias(const or_grid *g, double *spec)
{
    if (g->dealias == OR_DEALIAS_NONE) return;
    cplx *A = (cplx *)spec;
    const int nzc = g->nz / 2 + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++)
                if (dealias_zero(g, lx, ly, lz)) A[cidx(g, lx, ly, lz)] = 0;
}

/* Satisfy_strong_reality_condition_in_Array (FFF_pencil_basic.cc:139-179),
 * single-process form (num_y_procs == 1). */
void or_strong_reality(const or_grid *g, double *spec)
{
    cplx *A = (cplx *)spec;
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    cplx *plane = (cplx *)malloc(sizeof(cplx) * (size_t)nx * ny);
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++) plane[(size_t)lx * ny + ly] = A[cidx(g, lx, ly, 0)];
    for (int lx = nx / 2 + 1; lx < nx; lx++) {
        int mkx = -get_kx(g, lx);
        for (int ly = 0; ly < ny; ly++) {
            if (get_ky(g, ly) != ny / 2) {
                int mky = -get_ky(g, ly);
                int iy = mky >= 0 ? mky : mky + ny;
                A[cidx(g, lx, ly, 0)] = conj(plane[(size_t)mkx * ny + iy]);
            }
        }
    }
    for (int ly = 0; ly < ny; ly++) {
        if (get_ky(g, ly) < 0) {
            int mky = -get_ky(g, ly);
            A[cidx(g, 0, ly, 0)] = conj(plane[(size_t)0 * ny + mky]);
        }
    }
    free(plane);
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++) A[cidx(g, lx, ly, nzc - 1)] = 0;
}

/* Satisfy_weak_reality_condition_in_Array (FFF_pencil_basic.cc:181-186) */
void or_weak_reality(const or_grid *g, double *spec)
{
    cplx *A = (cplx *)spec;
    const int nzc = g->nz / 2 + 1;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++) A[cidx(g, lx, ly, nzc - 1)] = 0;
}

/* ------------------------------------------------------------------------ */
/* Nonlinear term                                                           */
/* ------------------------------------------------------------------------ */

static size_t nreal(const or_grid *g) { return (size_t)g->nx * g->ny * (size_t)(g->nz + 2); }
static size_t ncplx(const or_grid *g) { return (size_t)g->nx * g->ny * (size_t)(g->nz / 2 + 1); }

/* ArrayOps::Real_space_multiply (ArrayOps.cc:49-63): C = A*B */
static void real_mult(const or_grid *g, const double *a, const double *b, double *c)
{
    const size_t n = nreal(g);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) c[i] = a[i] * b[i];
}

/* B (op)= i K_d A for axis d (FFF_pencil_tr.cc:70-145); add != 0 adds. */
static void deriv(const or_grid *g, int d, const cplx *A, cplx *B, int add)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K = d == 0 ? get_kx(g, lx) * g->kfactor[0]
                         : d == 1 ? get_ky(g, ly) * g->kfactor[1]
                                  : lz * g->kfactor[2];
                size_t i = cidx(g, lx, ly, lz);
                cplx v = (I * K) * A[i];
                if (add) B[i] += v; else B[i] = v;
            }
}

/* Nlin_incompress::Compute_nlin(FluidVF&) (compute_nlin.cc:58-109) with
 * Compute_nlin_diag (compute_diag_offdiag.cc:47-65) and Compute_nlin_offdiag
 * (:68-112).  N_i = sum_j i K_j FT[u_j u_i].  Dealiases V in place (:61-62). */
void or_compute_nlin(const or_grid *g, double *V[3], double *N[3])
{
    const size_t nr = nreal(g);
    for (int c = 0; c < 3; c++) or_dealias(g, V[c]);
    double *u[3], *p = (double *)malloc(sizeof(double) * nr);
    for (int c = 0; c < 3; c++) {
        u[c] = (double *)malloc(sizeof(double) * nr);
        or_inverse(g, V[c], u[c]);
    }
    cplx *X = (cplx *)malloc(sizeof(cplx) * ncplx(g));
    cplx *n1 = (cplx *)N[0], *n2 = (cplx *)N[1], *n3 = (cplx *)N[2];
    /* diag */
    real_mult(g, u[0], u[0], p); or_forward(g, p, N[0]); deriv(g, 0, n1, n1, 0);
    real_mult(g, u[1], u[1], p); or_forward(g, p, N[1]); deriv(g, 1, n2, n2, 0);
    real_mult(g, u[2], u[2], p); or_forward(g, p, N[2]); deriv(g, 2, n3, n3, 0);
    /* offdiag */
    real_mult(g, u[0], u[1], p); or_forward(g, p, (double *)X);
    deriv(g, 1, X, n1, 1); deriv(g, 0, X, n2, 1);
    real_mult(g, u[1], u[2], p); or_forward(g, p, (double *)X);
    deriv(g, 2, X, n2, 1); deriv(g, 1, X, n3, 1);
    real_mult(g, u[2], u[0], p); or_forward(g, p, (double *)X);
    deriv(g, 2, X, n1, 1); deriv(g, 0, X, n3, 1);
    free(X); free(p);
    for (int c = 0; c < 3; c++) free(u[c]);
}

/* Compute_nlin(FluidVF& U, PlainFluidVF& W) (compute_nlin_ET_VF.cc:71-87) with
 * the (U,Plain) diag/offdiag (compute_diag_offdiag.cc:329-414):
 * N_i = sum_j i K_j FT[u_j w_i].  Dealiases V and W in place. */
void or_compute_nlin_giver(const or_grid *g, double *V[3], double *W[3], double *N[3])
{
    const size_t nr = nreal(g);
    for (int c = 0; c < 3; c++) { or_dealias(g, V[c]); or_dealias(g, W[c]); }
    double *u[3], *w[3], *p = (double *)malloc(sizeof(double) * nr);
    for (int c = 0; c < 3; c++) {
        u[c] = (double *)malloc(sizeof(double) * nr);
        w[c] = (double *)malloc(sizeof(double) * nr);
        or_inverse(g, V[c], u[c]);
        or_inverse(g, W[c], w[c]);
    }
    cplx *X = (cplx *)malloc(sizeof(cplx) * ncplx(g));
    cplx *n1 = (cplx *)N[0], *n2 = (cplx *)N[1], *n3 = (cplx *)N[2];
    real_mult(g, u[0], w[0], p); or_forward(g, p, N[0]); deriv(g, 0, n1, n1, 0);
    real_mult(g, u[1], w[1], p); or_forward(g, p, N[1]); deriv(g, 1, n2, n2, 0);
    real_mult(g, u[2], w[2], p); or_forward(g, p, N[2]); deriv(g, 2, n3, n3, 0);
    real_mult(g, u[0], w[1], p); or_forward(g, p, (double *)X); deriv(g, 0, X, n2, 1);
    real_mult(g, w[0], u[1], p); or_forward(g, p, (double *)X); deriv(g, 1, X, n1, 1);
    real_mult(g, u[1], w[2], p); or_forward(g, p, (double *)X); deriv(g, 1, X, n3, 1);
    real_mult(g, w[1], u[2], p); or_forward(g, p, (double *)X); deriv(g, 2, X, n2, 1);
    real_mult(g, u[2], w[0], p); or_forward(g, p, (double *)X); deriv(g, 2, X, n1, 1);
    real_mult(g, w[2], u[0], p); or_forward(g, p, (double *)X); deriv(g, 0, X, n3, 1);
    free(X); free(p);
    for (int c = 0; c < 3; c++) { free(u[c]); free(w[c]); }
}

/* Constant energy supply alpha-forcing F = alpha V on the band (r_in, r_out]
 * with alpha = eps / (2 E_band) -- the helicity-free limit of
 * FORCE::Compute_force_Carati_scheme_energy_supply (compute_force_Carati.cc:
 * 62-149) with one supply shell and Const_energy_supply_alpha_beta
 * (force_alpha_beta.cc:56-94, beta = 0).  E_band uses the energy weights of
 * Get_local_energy (FFF_pencil_energy.cc:59-67). */
static inline double uniform01(uint64_t seed, uint64_t ctr);

/* Universal::Compute_modes_in_shell_array (universal_basic.cc:104-143): the
 * number of modes of the full box, kx in (-Nx/2, Nx/2] (likewise ky, kz),
 * with ceil(K) = i, for i = 0..imax (only |k_d| <= imax / kfactor_d can
 * reach those shells). */
void or_modes_in_shell(const or_grid *g, int imax, double *count)
{
    for (int i = 0; i <= imax; i++) count[i] = 0;
    const int n[3] = {g->nx, g->ny, g->nz};
    int lim[3];
    for (int d = 0; d < 3; d++) {
        lim[d] = (int)ceil(imax / g->kfactor[d]) + 1;
        if (lim[d] > n[d] / 2) lim[d] = n[d] / 2;
    }
    for (int kx = -lim[0]; kx <= lim[0]; kx++)
        for (int ky = -lim[1]; ky <= lim[1]; ky++)
            for (int kz = -lim[2]; kz <= lim[2]; kz++) {
                if (kx <= -n[0] / 2 || ky <= -n[1] / 2 || kz <= -n[2] / 2) continue;
                const double a = kx * g->kfactor[0], b = ky * g->kfactor[1], c = kz * g->kfactor[2];
                const int idx = (int)ceil(sqrt(a * a + b * b + c * c));
                if (idx <= imax) count[idx] += 1;
            }
}

/* Random band forcing: Compute_force_using_random_noise_basic
 * (compute_force_random.cc:65-130) with Put_or_add_force_vector
 * (force_amp_phase.cc:119-170), Helical_to_Craya and Craya_to_cartesian
 * (FFF_pencil_inline.h:234-248) about the z axis (anisotropy_dirn 3).  On
 * every stored mode with r_in < K <= r_out: e_k = supply_e(i) /
 * modes_in_shell(i), h_k likewise, i = ceil(K); F+- = sqrt((e_k +- h_k/K) /
 * tau) exp(i phi+-); U1 = i (F- - F+), U2 = F- + F+; F = (vh1, vh2, vpll)
 * with vpll = -U2 sin(theta), vh1 = U1 sin(phi) + U2 cos(theta) cos(phi),
 * vh2 = -U1 cos(phi) + U2 sin(phi) cos(theta).  The supply spectra, which
 * the reference reads from a file, are here the rates force_eps and
 * force_helicity split evenly over the shells i = floor(r_in)+1 ..
 * ceil(r_out); the phases phi+- = 2 pi U(seed, 2 m), 2 pi U(seed, 2 m + 1),
 * m the mode's index in the (Nx, Ny, Nz/2+1) layout, from the counter-based
 * generator of the seeded initial field (the reference's Blitz generator
 * is seeded by my_id * time(0), compute_force_random.cc:143-167). */
static void random_force(const or_grid *g, const or_fluid *f, double *F[3])
{
    const int nzc = g->nz / 2 + 1;
    const int imax = (int)ceil(f->force_r_out), i0 = (int)floor(f->force_r_in) + 1;
    const int nsh = imax - i0 + 1;
    if (nsh < 1) return;
    double *cnt = (double *)calloc((size_t)imax + 1, sizeof(double));
    or_modes_in_shell(g, imax, cnt);
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const double K = or_kmag(g, lx, ly, lz);
                if (!(K > f->force_r_in && K <= f->force_r_out)) continue;
                const int i = (int)ceil(K);
                const double e = f->force_eps / nsh / cnt[i], h = f->force_helicity / nsh / cnt[i];
                const double ap = sqrt((e + h / K) / f->force_tau), am = sqrt((e - h / K) / f->force_tau);
                const size_t m = cidx(g, lx, ly, lz);
                const double php = 2 * M_PI * uniform01(f->force_seed, 2 * (uint64_t)m);
                const double phm = 2 * M_PI * uniform01(f->force_seed, 2 * (uint64_t)m + 1);
                const cplx Fp = ap * (cos(php) + I * sin(php)), Fm = am * (cos(phm) + I * sin(phm));
                const cplx U1 = I * (Fm - Fp), U2 = Fm + Fp;
                const double th = or_polar_angle(g, lx, ly, lz, 3), ph = or_azimuthal_angle(g, lx, ly, lz, 3);
                ((cplx *)F[0])[m] = U1 * sin(ph) + U2 * cos(th) * cos(ph);
                ((cplx *)F[1])[m] = -U1 * cos(ph) + U2 * sin(ph) * cos(th);
                ((cplx *)F[2])[m] = -U2 * sin(th);
            }
    free(cnt);
}

void or_compute_force(const or_grid *g, const or_fluid *f, double *const V[3], double *F[3])
{
    const size_t nc = ncplx(g);
    for (int c = 0; c < 3; c++) memset(F[c], 0, sizeof(cplx) * nc);
    if (f->force_kind == 0) return;
    if (f->force_kind == 2) {
        random_force(g, f, F);
        return;
    }
    const int nzc = g->nz / 2 + 1;
    double eband = 0.0;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K = or_kmag(g, lx, ly, lz);
                if (K > f->force_r_in && K <= f->force_r_out) {
                    double w = lz == 0 ? 0.5 : 1.0;
                    size_t i = cidx(g, lx, ly, lz);
                    for (int c = 0; c < 3; c++) {
                        cplx v = ((cplx *)V[c])[i];
                        eband += w * (creal(v) * creal(v) + cimag(v) * cimag(v));
                    }
                }
            }
    if (!(eband > 0.0)) return;
    double alpha = f->force_eps / (2.0 * eband);
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K = or_kmag(g, lx, ly, lz);
                if (K > f->force_r_in && K <= f->force_r_out) {
                    size_t i = cidx(g, lx, ly, lz);
                    for (int c = 0; c < 3; c++) ((cplx *)F[c])[i] = alpha * ((cplx *)V[c])[i];
                }
            }
}

/* Add_force (Tadvance_functions.cc:45-53), Pressure::Compute_pressure
 * (Pressure.cc:56-72: divergence universal_basic.cc:343-363, divide_ksqr
 * :202-225), Compute_rhs (Tadvance_functions.cc:148-191):
 * N <- -(N - F + i K p), p = i K.(N - F) / K^2, p(0) = 0. */
void or_compute_rhs(const or_grid *g, double *N[3], double *const F[3])
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    cplx *n1 = (cplx *)N[0], *n2 = (cplx *)N[1], *n3 = (cplx *)N[2];
    const cplx *f1 = F ? (const cplx *)F[0] : NULL, *f2 = F ? (const cplx *)F[1] : NULL,
               *f3 = F ? (const cplx *)F[2] : NULL;
#pragma omp parallel for collapse(2) schedule(static)
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                size_t i = cidx(g, lx, ly, lz);
                double kx = get_kx(g, lx) * g->kfactor[0];
                double ky = get_ky(g, ly) * g->kfactor[1];
                double kz = lz * g->kfactor[2];
                cplx a = n1[i], b = n2[i], c = n3[i];
                if (f1) { a -= f1[i]; b -= f2[i]; c -= f3[i]; }
                cplx div = (I * kx) * a;
                div += (I * ky) * b;
                div += (I * kz) * c;
                double ks = ksqr(g, lx, ly, lz);
                cplx p = (lx == 0 && ly == 0 && lz == 0) ? 0 : div / ks;
                a += (I * kx) * p;
                b += (I * ky) * p;
                c += (I * kz) * p;
                n1[i] = -a; n2[i] = -b; n3[i] = -c;
            }
}

/* The exponent of Array_exp_ksqr: factor K^2 (universal_basic.cc:235-255),
 * or with hyperviscosity (factor + hyper_factor K^(q-2)) K^2
 * (universal_basic.cc:265-291), K^(q-2) = (K^2)^((q-2)/2) with integer
 * division (q = 4: K^2), the power by repeated multiplication; q = 0: none. */
typedef struct { double f, fh; int q; } lfac;

static double lexpo(const lfac *L, double ks)
{
    if (L->q == 0) return L->f * ks;
    const int n = L->q == 4 ? 1 : (L->q - 2) / 2;
    double kp = 1.0;
    for (int i = 0; i < n; i++) kp *= ks;
    return (L->f + L->fh * kp) * ks;
}

/* the factors of a fluid for a multiple a of dt: (-nu a dt, -nu_hyper a dt, q) */
static lfac fluid_lfac(const or_fluid *f, double a)
{
    lfac L = {-f->nu * a * f->dt, -f->nu_hyper * a * f->dt, f->hyper_exponent};
    return L;
}

/* Array_exp_ksqr: A *= exp(exponent) */
static void exp_ksqr_l(const or_grid *g, double *spec, const lfac *L)
{
    cplx *A = (cplx *)spec;
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) A[cidx(g, lx, ly, lz)] *= exp(lexpo(L, ksqr(g, lx, ly, lz)));
}

/* Array_exp_ksqr (universal_basic.cc:235-255): A *= exp(factor*K^2) */
static void exp_ksqr(const or_grid *g, double *spec, double factor)
{
    lfac L = {factor, 0.0, 0};
    exp_ksqr_l(g, spec, &L);
}

/* FluidVF::Mult_field_exp_ksqr_dt / Mult_nlin_exp_ksqr_dt (FluidVF.cc:339-389) */
static void mult_exp(const or_grid *g, const or_fluid *f, double *A[3], double a)
{
    if (fabs(a) > 1e-15) {
        lfac L = fluid_lfac(f, a);
        for (int c = 0; c < 3; c++) exp_ksqr_l(g, A[c], &L);
    }
}

/* FluidVF::Add_nlin_factor_dt (FluidVF.cc:308-316): Y += (factor*dt)*N */
static void add_factor(const or_grid *g, double *Y[3], double *const N[3], double factor, double dt)
{
    if (fabs(factor) <= 1e-5) return;
    const size_t nc = ncplx(g);
    const double s = factor * dt;
    for (int c = 0; c < 3; c++) {
        cplx *y = (cplx *)Y[c];
        const cplx *n = (const cplx *)N[c];
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < nc; i++) y[i] += s * n[i];
    }
}

/* Compute_force_TO_rhs (Tadvance_functions.cc:441-453) */
static void force_to_rhs(const or_grid *g, const or_fluid *f, double *V[3], double *N[3], double *F[3])
{
    or_compute_force(g, f, V, F);
    or_compute_nlin(g, V, N);
    or_compute_rhs(g, N, F);
}

/* Single_time_step (Tadvance_functions.cc:298-324) */
static void single_time_step(const or_grid *g, const or_fluid *f, double *V[3], double *N[3],
                             double a, double b, double c)
{
    const double eps = 1e-15;
    if (fabs(a) < eps && fabs(b) < eps) {
        add_factor(g, V, N, c, f->dt);
    } else if (fabs(a) < eps && fabs(b) > eps) {
        mult_exp(g, f, N, b);
        add_factor(g, V, N, c, f->dt);
    } else if (fabs(b) < eps && fabs(a) > eps) {
        mult_exp(g, f, V, a);
        add_factor(g, V, N, c, f->dt);
    } else if (fabs(a - b) < eps && fabs(a) > eps) {
        add_factor(g, V, N, c, f->dt);
        mult_exp(g, f, V, a);
    } else if (fabs(a) > eps && fabs(b) > eps) {
        mult_exp(g, f, V, a - b);
        add_factor(g, V, N, c, f->dt);
        mult_exp(g, f, V, b);
    }
}

/* Compute_RK4_parts (Tadvance_functions.cc:566-570) */
static void rk4_parts(const or_grid *g, const or_fluid *f, double *N[3], double *acc[3], double b,
                      double factor)
{
    mult_exp(g, f, N, b);
    add_factor(g, acc, N, factor, f->dt);
}

/* Time_advance_incompress::RK4 (RK.cc:91-170, LES branch omitted) followed
 * by the strong reality condition (Time_advance.cc:60-61). */
void or_rk4_step(const or_grid *g, const or_fluid *f, double *V[3])
{
    const size_t nc = ncplx(g);
    double *V0[3], *acc[3], *N[3], *F[3];
    for (int c = 0; c < 3; c++) {
        V0[c] = (double *)malloc(sizeof(cplx) * nc);
        acc[c] = (double *)calloc(nc, sizeof(cplx));
        N[c] = (double *)malloc(sizeof(cplx) * nc);
        F[c] = (double *)malloc(sizeof(cplx) * nc);
        memcpy(V0[c], V[c], sizeof(cplx) * nc);
    }
    /* step 1 */
    force_to_rhs(g, f, V, N, F);
    single_time_step(g, f, V, N, 0.5, 0.5, 0.5);
    rk4_parts(g, f, N, acc, 1.0, 1.0 / 6);
    /* step 2 */
    force_to_rhs(g, f, V, N, F);
    for (int c = 0; c < 3; c++) memcpy(V[c], V0[c], sizeof(cplx) * nc);
    single_time_step(g, f, V, N, 0.5, 0, 0.5);
    rk4_parts(g, f, N, acc, 0.5, 1.0 / 3);
    /* step 3 */
    force_to_rhs(g, f, V, N, F);
    for (int c = 0; c < 3; c++) memcpy(V[c], V0[c], sizeof(cplx) * nc);
    single_time_step(g, f, V, N, 1, 0.5, 1);
    rk4_parts(g, f, N, acc, 0.5, 1.0 / 3);
    /* step 4 */
    force_to_rhs(g, f, V, N, F);
    rk4_parts(g, f, N, acc, 0, 1.0 / 6);
    /* final */
    for (int c = 0; c < 3; c++) memcpy(V[c], V0[c], sizeof(cplx) * nc);
    mult_exp(g, f, V, 1.0);
    for (int c = 0; c < 3; c++) {
        cplx *v = (cplx *)V[c];
        const cplx *a = (const cplx *)acc[c];
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < nc; i++) v[i] = v[i] + a[i];
    }
    if (g->strong_reality)
        for (int c = 0; c < 3; c++) or_strong_reality(g, V[c]);
    for (int c = 0; c < 3; c++) { free(V0[c]); free(acc[c]); free(N[c]); free(F[c]); }
}

/* Euler (RK.cc:61-66) */
void or_euler_step(const or_grid *g, const or_fluid *f, double *V[3])
{
    const size_t nc = ncplx(g);
    double *N[3], *F[3];
    for (int c = 0; c < 3; c++) {
        N[c] = (double *)malloc(sizeof(cplx) * nc);
        F[c] = (double *)malloc(sizeof(cplx) * nc);
    }
    force_to_rhs(g, f, V, N, F);
    single_time_step(g, f, V, N, 1, 1, 1);
    if (g->strong_reality)
        for (int c = 0; c < 3; c++) or_strong_reality(g, V[c]);
    for (int c = 0; c < 3; c++) { free(N[c]); free(F[c]); }
}

/* ------------------------------------------------------------------------ */
/* MHD (Elsasser form)                                                       */
/* ------------------------------------------------------------------------ */

/* Nlin_incompress::Compute_nlin(FluidVF& U, FluidVF& W) (compute_nlin_VF_MHD.cc:71-123):
 * dealias U, W in place; inverse; Z+ = u + b, Z- = u - b in real space
 * (MHD::UB_to_Elsasser_real_field, MHD.cc:75-88); N+_i = sum_j d_j(Z-_j Z+_i)
 * into NU, N-_i = sum_j d_j(Z+_j Z-_i) into NW, with the diag/offdiag order
 * of compute_diag_offdiag.cc:139-257; then NU = (N+ + N-)/2, NW = (N+ - N-)/2
 * (MHD::Elsasser_to_UB_nlin, MHD.cc:150-163). */
void or_compute_nlin_mhd(const or_grid *g, double *V[3], double *W[3], double *NU[3], double *NW[3])
{
    const size_t nr = nreal(g), nc = ncplx(g);
    for (int c = 0; c < 3; c++) { or_dealias(g, V[c]); or_dealias(g, W[c]); }
    double *zp[3], *zm[3], *p = (double *)malloc(sizeof(double) * nr);
    for (int c = 0; c < 3; c++) {
        zp[c] = (double *)malloc(sizeof(double) * nr);
        zm[c] = (double *)malloc(sizeof(double) * nr);
        or_inverse(g, V[c], zp[c]);
        or_inverse(g, W[c], zm[c]);
        for (size_t i = 0; i < nr; i++) {
            double u = zp[c][i], b = zm[c][i];
            zp[c][i] = u + b;
            zm[c][i] = u - b;
        }
    }
    cplx *X = (cplx *)malloc(sizeof(cplx) * nc);
    cplx *u1 = (cplx *)NU[0], *u2 = (cplx *)NU[1], *u3 = (cplx *)NU[2];
    cplx *w1 = (cplx *)NW[0], *w2 = (cplx *)NW[1], *w3 = (cplx *)NW[2];
    /* diag */
    real_mult(g, zp[0], zm[0], p); or_forward(g, p, NU[0]); deriv(g, 0, u1, u1, 0);
    real_mult(g, zp[1], zm[1], p); or_forward(g, p, NU[1]); deriv(g, 1, u2, u2, 0);
    real_mult(g, zp[2], zm[2], p); or_forward(g, p, NU[2]); deriv(g, 2, u3, u3, 0);
    memcpy(w1, u1, sizeof(cplx) * nc); memcpy(w2, u2, sizeof(cplx) * nc); memcpy(w3, u3, sizeof(cplx) * nc);
    /* offdiag */
    real_mult(g, zp[0], zm[1], p); or_forward(g, p, (double *)X); deriv(g, 1, X, u1, 1); deriv(g, 0, X, w2, 1);
    real_mult(g, zm[0], zp[1], p); or_forward(g, p, (double *)X); deriv(g, 1, X, w1, 1); deriv(g, 0, X, u2, 1);
    real_mult(g, zp[1], zm[2], p); or_forward(g, p, (double *)X); deriv(g, 2, X, u2, 1); deriv(g, 1, X, w3, 1);
    real_mult(g, zm[1], zp[2], p); or_forward(g, p, (double *)X); deriv(g, 2, X, w2, 1); deriv(g, 1, X, u3, 1);
    real_mult(g, zp[2], zm[0], p); or_forward(g, p, (double *)X); deriv(g, 2, X, w1, 1); deriv(g, 0, X, u3, 1);
    real_mult(g, zm[2], zp[0], p); or_forward(g, p, (double *)X); deriv(g, 2, X, u1, 1); deriv(g, 0, X, w3, 1);
    for (int c = 0; c < 3; c++) {
        cplx *a = (cplx *)NU[c], *b = (cplx *)NW[c];
        for (size_t i = 0; i < nc; i++) {
            cplx t = a[i];
            a[i] = 0.5 * (a[i] + b[i]);
            b[i] = 0.5 * (t - b[i]);
        }
    }
    free(X); free(p);
    for (int c = 0; c < 3; c++) { free(zp[c]); free(zm[c]); }
}

/* Compute_rhs(U, W, P) (Tadvance_functions.cc:262-270): U as the fluid, W
 * negated (no pressure, no forcing on W). */
static void mhd_force_to_rhs(const or_grid *g, const or_fluid *fu, double *V[3], double *W[3], double *NU[3],
                             double *NW[3], double *F[3])
{
    or_compute_force(g, fu, V, F);
    or_compute_nlin_mhd(g, V, W, NU, NW);
    or_compute_rhs(g, NU, F);
    const size_t nc = ncplx(g);
    for (int c = 0; c < 3; c++) {
        cplx *b = (cplx *)NW[c];
        for (size_t i = 0; i < nc; i++) b[i] = -b[i];
    }
}

/* Time_advance_incompress::RK4(U, W, P, F) (RK_VF_MHD.cc:126-190): the same
 * Lawson stages for both fields, U with nu = fu->nu, W with eta = eta. */
void or_rk4_step_mhd(const or_grid *g, const or_fluid *fu, double eta, double *V[3], double *W[3])
{
    const size_t nc = ncplx(g);
    or_fluid fb = *fu;
    fb.nu = eta;
    fb.force_kind = 0;
    double *V0[3], *W0[3], *accV[3], *accW[3], *NU[3], *NW[3], *F[3];
    for (int c = 0; c < 3; c++) {
        V0[c] = (double *)malloc(sizeof(cplx) * nc);
        W0[c] = (double *)malloc(sizeof(cplx) * nc);
        accV[c] = (double *)calloc(nc, sizeof(cplx));
        accW[c] = (double *)calloc(nc, sizeof(cplx));
        NU[c] = (double *)malloc(sizeof(cplx) * nc);
        NW[c] = (double *)malloc(sizeof(cplx) * nc);
        F[c] = (double *)malloc(sizeof(cplx) * nc);
        memcpy(V0[c], V[c], sizeof(cplx) * nc);
        memcpy(W0[c], W[c], sizeof(cplx) * nc);
    }
    mhd_force_to_rhs(g, fu, V, W, NU, NW, F);
    single_time_step(g, fu, V, NU, 0.5, 0.5, 0.5);
    single_time_step(g, &fb, W, NW, 0.5, 0.5, 0.5);
    rk4_parts(g, fu, NU, accV, 1.0, 1.0 / 6);
    rk4_parts(g, &fb, NW, accW, 1.0, 1.0 / 6);

    mhd_force_to_rhs(g, fu, V, W, NU, NW, F);
    for (int c = 0; c < 3; c++) { memcpy(V[c], V0[c], sizeof(cplx) * nc); memcpy(W[c], W0[c], sizeof(cplx) * nc); }
    single_time_step(g, fu, V, NU, 0.5, 0, 0.5);
    single_time_step(g, &fb, W, NW, 0.5, 0, 0.5);
    rk4_parts(g, fu, NU, accV, 0.5, 1.0 / 3);
    rk4_parts(g, &fb, NW, accW, 0.5, 1.0 / 3);

    mhd_force_to_rhs(g, fu, V, W, NU, NW, F);
    for (int c = 0; c < 3; c++) { memcpy(V[c], V0[c], sizeof(cplx) * nc); memcpy(W[c], W0[c], sizeof(cplx) * nc); }
    single_time_step(g, fu, V, NU, 1, 0.5, 1);
    single_time_step(g, &fb, W, NW, 1, 0.5, 1);
    rk4_parts(g, fu, NU, accV, 0.5, 1.0 / 3);
    rk4_parts(g, &fb, NW, accW, 0.5, 1.0 / 3);

    mhd_force_to_rhs(g, fu, V, W, NU, NW, F);
    rk4_parts(g, fu, NU, accV, 0, 1.0 / 6);
    rk4_parts(g, &fb, NW, accW, 0, 1.0 / 6);

    for (int c = 0; c < 3; c++) { memcpy(V[c], V0[c], sizeof(cplx) * nc); memcpy(W[c], W0[c], sizeof(cplx) * nc); }
    mult_exp(g, fu, V, 1.0);
    mult_exp(g, &fb, W, 1.0);
    for (int c = 0; c < 3; c++) {
        cplx *v = (cplx *)V[c], *w = (cplx *)W[c];
        const cplx *a = (const cplx *)accV[c], *b = (const cplx *)accW[c];
        for (size_t i = 0; i < nc; i++) { v[i] = v[i] + a[i]; w[i] = w[i] + b[i]; }
    }
    if (g->strong_reality)
        for (int c = 0; c < 3; c++) { or_strong_reality(g, V[c]); or_strong_reality(g, W[c]); }
    for (int c = 0; c < 3; c++) {
        free(V0[c]); free(W0[c]); free(accV[c]); free(accW[c]); free(NU[c]); free(NW[c]); free(F[c]);
    }
}

/* Get_local_energy(A, B) (FFF_pencil_energy.cc:75-84):
 * sum Re(A B*) - sum_{kz=0} Re(A B*)/2; cross helicity = sum over components
 * (Correlation::Get_cross_helicity, Correlation.cc:225-230). */
double or_cross_energy(const or_grid *g, const double *a, const double *b)
{
    const cplx *A = (const cplx *)a, *B = (const cplx *)b;
    const size_t nc = ncplx(g);
    const int nzc = g->nz / 2 + 1;
    double s = 0.0, s0 = 0.0;
    for (size_t i = 0; i < nc; i++) s += creal(A[i]) * creal(B[i]) + cimag(A[i]) * cimag(B[i]);
    for (size_t i = 0; i < nc; i += nzc) s0 += creal(A[i]) * creal(B[i]) + cimag(A[i]) * cimag(B[i]);
    return s - s0 / 2;
}

/* EnergyTr::Compute_flux(FluidVF& U, FluidVF& W) (compute_flux.cc:158-211),
 * the u< -> b> and b< -> b> parts:
 *   u2b[s] = sum_{K > r_s} 2m Re(N.b*), N = FT[d_j(b_j u<_i)]
 *   b2b[s] = -sum_{K > r_s} 2m Re(N.b*), N = FT[d_j(u_j b<_i)]. */
void or_flux_mhd(const or_grid *g, double *V[3], double *W[3], const double *radii, int nspheres, double *u2b,
                 double *b2b)
{
    const size_t nc = ncplx(g);
    double *G[3], *N[3];
    for (int c = 0; c < 3; c++) {
        G[c] = (double *)malloc(sizeof(cplx) * nc);
        N[c] = (double *)malloc(sizeof(cplx) * nc);
    }
    u2b[0] = b2b[0] = 0.0;
    for (int s = 1; s <= nspheres; s++) {
        for (int c = 0; c < 3; c++) or_fill_array_shell(g, V[c], G[c], 0.0, radii[s]);
        or_compute_nlin_giver(g, W, G, N);
        double p = 0.0;
        for (int c = 0; c < 3; c++) p += or_shell_mult_single(g, N[c], W[c], radii[s], INFINITY);
        u2b[s] = p;
    }
    for (int s = 1; s <= nspheres; s++) {
        for (int c = 0; c < 3; c++) or_fill_array_shell(g, W[c], G[c], 0.0, radii[s]);
        or_compute_nlin_giver(g, V, G, N);
        double p = 0.0;
        for (int c = 0; c < 3; c++) p += or_shell_mult_single(g, N[c], W[c], radii[s], INFINITY);
        b2b[s] = -p;
    }
    for (int c = 0; c < 3; c++) { free(G[c]); free(N[c]); }
}

/* ------------------------------------------------------------------------ */
/* Tangent and adjoint of the NS right-hand side and of the Lawson RK4 step  */
/* (no counterpart in the reference; this build's FP64 CPU adjoint)          */
/* ------------------------------------------------------------------------ */
/*
 * f(V) = -P NL(D V), NL(U)_i = sum_j i K_j F[I(U_j) I(U_i)] (or_compute_nlin +
 * or_compute_rhs without forcing).  Inner product: the real dot product of
 * the stored half-spectrum coefficients (Re, Im).  With c = 1 on the kz = 0
 * and kz = Nz/2 planes and 2 elsewhere, the transposes of the transforms are
 * I^T = N^3 diag(c) F and F^T = N^-3 I diag(1/c); (i K)^T = -i K; P and D are
 * symmetric.  Hence
 *   J^T L = -D diag(c) F[ sum_j u_j I(diag(1/c)(-i K_m M_j - i K_j M_m)) ]_m,
 *   M = P L, u = I(D V).
 */

static double cweight(const or_grid *g, int lz) { return (lz == 0 || lz == g->nz / 2) ? 1.0 : 2.0; }

static void project(const or_grid *g, cplx *a0, cplx *a1, cplx *a2)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                size_t i = cidx(g, lx, ly, lz);
                double kx = get_kx(g, lx) * g->kfactor[0], ky = get_ky(g, ly) * g->kfactor[1], kz = lz * g->kfactor[2];
                double ks = ksqr(g, lx, ly, lz);
                if (ks == 0.0) continue;
                cplx d = kx * a0[i] + ky * a1[i] + kz * a2[i];
                a0[i] -= kx * d / ks;
                a1[i] -= ky * d / ks;
                a2[i] -= kz * d / ks;
            }
}

/* out = J(V) dV (tangent of the right-hand side) */
void or_rhs_jvp(const or_grid *g, double *const V[3], double *const dV[3], double *out[3])
{
    const size_t nr = nreal(g), nc = ncplx(g);
    double *u[3], *du[3], *p = (double *)malloc(sizeof(double) * nr);
    double *tmp = (double *)malloc(sizeof(cplx) * nc);
    for (int c = 0; c < 3; c++) {
        u[c] = (double *)malloc(sizeof(double) * nr);
        du[c] = (double *)malloc(sizeof(double) * nr);
        memcpy(tmp, V[c], sizeof(cplx) * nc); or_dealias(g, tmp); or_inverse(g, tmp, u[c]);
        memcpy(tmp, dV[c], sizeof(cplx) * nc); or_dealias(g, tmp); or_inverse(g, tmp, du[c]);
        memset(out[c], 0, sizeof(cplx) * nc);
    }
    cplx *X = (cplx *)malloc(sizeof(cplx) * nc);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < nr; q++) p[q] = du[j][q] * u[i][q] + u[j][q] * du[i][q];
            or_forward(g, p, (double *)X);
            deriv(g, j, X, (cplx *)out[i], 1);
        }
    project(g, (cplx *)out[0], (cplx *)out[1], (cplx *)out[2]);
    for (int c = 0; c < 3; c++) {
        cplx *o = (cplx *)out[c];
        for (size_t q = 0; q < nc; q++) o[q] = -o[q];
    }
    free(X); free(p); free(tmp);
    for (int c = 0; c < 3; c++) { free(u[c]); free(du[c]); }
}

/* out = J(V)^T L (adjoint of the right-hand side) */
void or_rhs_vjp(const or_grid *g, double *const V[3], double *const Lam[3], double *out[3])
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    const size_t nr = nreal(g), nc = ncplx(g);
    cplx *M[3];
    double *u[3], *sig[6], *q = (double *)malloc(sizeof(double) * nr);
    double *tmp = (double *)malloc(sizeof(cplx) * nc);
    for (int c = 0; c < 3; c++) {
        M[c] = (cplx *)malloc(sizeof(cplx) * nc);
        memcpy(M[c], Lam[c], sizeof(cplx) * nc);
        u[c] = (double *)malloc(sizeof(double) * nr);
        memcpy(tmp, V[c], sizeof(cplx) * nc); or_dealias(g, tmp); or_inverse(g, tmp, u[c]);
    }
    project(g, M[0], M[1], M[2]);
    /* sigma_jm = I(diag(1/c)(-i K_m M_j - i K_j M_m)), j <= m */
    int k = 0;
    int jj[6], mm[6];
    for (int j = 0; j < 3; j++)
        for (int m = j; m < 3; m++) {
            jj[k] = j; mm[k] = m;
            cplx *S = (cplx *)tmp;
#pragma omp parallel for collapse(2) schedule(static)
            for (int lx = 0; lx < nx; lx++)
                for (int ly = 0; ly < ny; ly++)
                    for (int lz = 0; lz < nzc; lz++) {
                        size_t i = cidx(g, lx, ly, lz);
                        double K[3] = {get_kx(g, lx) * g->kfactor[0], get_ky(g, ly) * g->kfactor[1], lz * g->kfactor[2]};
                        S[i] = (-I * K[m] * M[j][i] - I * K[j] * M[m][i]) / cweight(g, lz);
                    }
            sig[k] = (double *)malloc(sizeof(double) * nr);
            or_inverse(g, tmp, sig[k]);
            k++;
        }
    for (int m = 0; m < 3; m++) {
#pragma omp parallel for schedule(static)
        for (size_t r = 0; r < nr; r++) {
            double acc = 0.0;
            for (int t = 0; t < 6; t++) {
                if (mm[t] == m) acc += u[jj[t]][r] * sig[t][r];
                else if (jj[t] == m) acc += u[mm[t]][r] * sig[t][r];
            }
            q[r] = acc;
        }
        or_forward(g, q, out[m]);
        cplx *o = (cplx *)out[m];
#pragma omp parallel for collapse(2) schedule(static)
        for (int lx = 0; lx < nx; lx++)
            for (int ly = 0; ly < ny; ly++)
                for (int lz = 0; lz < nzc; lz++) {
                    size_t i = cidx(g, lx, ly, lz);
                    o[i] = dealias_zero(g, lx, ly, lz) ? 0 : -cweight(g, lz) * o[i];
                }
    }
    free(q); free(tmp);
    for (int c = 0; c < 3; c++) { free(M[c]); free(u[c]); }
    for (int t = 0; t < 6; t++) free(sig[t]);
}

static void scale_add(const or_grid *g, cplx *y, const cplx *x, double a, const lfac *fac)
{
    /* y += a * exp(exponent of fac) * x, fac = NULL for 1 */
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                size_t i = cidx(g, lx, ly, lz);
                double f = fac ? exp(lexpo(fac, ksqr(g, lx, ly, lz))) : 1.0;
                y[i] += a * f * x[i];
            }
}

/* Adjoint of one Lawson RK4 step without forcing and without the reality
 * condition: lam <- (dV_{n+1}/dV_n)^T lam, linearised at V (= V_n). */
void or_rk4_step_vjp(const or_grid *g, const or_fluid *f, double *const V[3], double *lam[3])
{
    const size_t nc = ncplx(g);
    const double dt = f->dt;
    const lfac fh = fluid_lfac(f, 0.5), f1 = fluid_lfac(f, 1.0);
    double *St[4][3], *G[3], *lg[3], *lv[3][3];
    /* stage inputs V0..V3 */
    for (int s = 0; s < 4; s++)
        for (int c = 0; c < 3; c++) St[s][c] = (double *)malloc(sizeof(cplx) * nc);
    for (int c = 0; c < 3; c++) {
        memcpy(St[0][c], V[c], sizeof(cplx) * nc);
        G[c] = (double *)malloc(sizeof(cplx) * nc);
        lg[c] = (double *)malloc(sizeof(cplx) * nc);
        for (int s = 0; s < 3; s++) lv[s][c] = (double *)calloc(nc, sizeof(cplx));
    }
    {
        double *W[3];
        for (int c = 0; c < 3; c++) { W[c] = (double *)malloc(sizeof(cplx) * nc); memcpy(W[c], V[c], sizeof(cplx) * nc); }
        /* V1 = Eh (D V0 + dt/2 f(V0)) */
        or_compute_nlin(g, W, G); or_compute_rhs(g, G, NULL);
        for (int c = 0; c < 3; c++) {
            cplx *v1 = (cplx *)St[1][c];
            memcpy(v1, W[c], sizeof(cplx) * nc);  /* W is D V0 after compute_nlin */
            scale_add(g, v1, (cplx *)G[c], 0.5 * dt, NULL);
            exp_ksqr_l(g, St[1][c], &fh);
        }
        /* V2 = Eh V0 + dt/2 f(V1) */
        for (int c = 0; c < 3; c++) memcpy(W[c], St[1][c], sizeof(cplx) * nc);
        or_compute_nlin(g, W, G); or_compute_rhs(g, G, NULL);
        for (int c = 0; c < 3; c++) {
            memcpy(St[2][c], V[c], sizeof(cplx) * nc);
            exp_ksqr_l(g, St[2][c], &fh);
            scale_add(g, (cplx *)St[2][c], (cplx *)G[c], 0.5 * dt, NULL);
        }
        /* V3 = Eh (Eh V0 + dt f(V2)) */
        for (int c = 0; c < 3; c++) memcpy(W[c], St[2][c], sizeof(cplx) * nc);
        or_compute_nlin(g, W, G); or_compute_rhs(g, G, NULL);
        for (int c = 0; c < 3; c++) {
            memcpy(St[3][c], V[c], sizeof(cplx) * nc);
            exp_ksqr_l(g, St[3][c], &fh);
            scale_add(g, (cplx *)St[3][c], (cplx *)G[c], dt, NULL);
            exp_ksqr_l(g, St[3][c], &fh);
        }
        for (int c = 0; c < 3; c++) free(W[c]);
    }
    /* backward */
    double *l4[3];
    for (int c = 0; c < 3; c++) { l4[c] = (double *)malloc(sizeof(cplx) * nc); memcpy(l4[c], lam[c], sizeof(cplx) * nc); }
    /* lam_V3 = dt/6 J3^T lam4 */
    for (int c = 0; c < 3; c++) { memcpy(lg[c], l4[c], sizeof(cplx) * nc); }
    or_rhs_vjp(g, St[3], lg, lv[2]);
    for (int c = 0; c < 3; c++) { cplx *x = (cplx *)lv[2][c]; for (size_t i = 0; i < nc; i++) x[i] *= dt / 6; }
    /* lam_g2 = dt/3 Eh lam4 + dt Eh lam_V3 ; lam_V2 = J2^T lam_g2 */
    for (int c = 0; c < 3; c++) {
        memset(lg[c], 0, sizeof(cplx) * nc);
        scale_add(g, (cplx *)lg[c], (cplx *)l4[c], dt / 3, &fh);
        scale_add(g, (cplx *)lg[c], (cplx *)lv[2][c], dt, &fh);
    }
    or_rhs_vjp(g, St[2], lg, lv[1]);
    /* lam_g1 = dt/3 Eh lam4 + dt/2 lam_V2 ; lam_V1 = J1^T lam_g1 */
    for (int c = 0; c < 3; c++) {
        memset(lg[c], 0, sizeof(cplx) * nc);
        scale_add(g, (cplx *)lg[c], (cplx *)l4[c], dt / 3, &fh);
        scale_add(g, (cplx *)lg[c], (cplx *)lv[1][c], dt / 2, NULL);
    }
    or_rhs_vjp(g, St[1], lg, lv[0]);
    /* lam_g0 = dt/6 E1 lam4 + dt/2 Eh lam_V1 */
    for (int c = 0; c < 3; c++) {
        memset(lg[c], 0, sizeof(cplx) * nc);
        scale_add(g, (cplx *)lg[c], (cplx *)l4[c], dt / 6, &f1);
        scale_add(g, (cplx *)lg[c], (cplx *)lv[0][c], dt / 2, &fh);
    }
    or_rhs_vjp(g, St[0], lg, G);
    /* lam_V0 = E1 lam4 + Eh^2 lam_V3 + Eh lam_V2 + D Eh lam_V1 + J0^T lam_g0 */
    for (int c = 0; c < 3; c++) {
        cplx *out = (cplx *)lam[c];
        memset(out, 0, sizeof(cplx) * nc);
        scale_add(g, out, (cplx *)l4[c], 1.0, &f1);
        const lfac f2 = {2 * fh.f, 2 * fh.fh, fh.q};
        scale_add(g, out, (cplx *)lv[2][c], 1.0, &f2);
        scale_add(g, out, (cplx *)lv[1][c], 1.0, &fh);
        double *t = (double *)malloc(sizeof(cplx) * nc);
        memcpy(t, lv[0][c], sizeof(cplx) * nc);
        or_dealias(g, t);
        scale_add(g, out, (cplx *)t, 1.0, &fh);
        free(t);
        scale_add(g, out, (cplx *)G[c], 1.0, NULL);
    }
    for (int s = 0; s < 4; s++)
        for (int c = 0; c < 3; c++) free(St[s][c]);
    for (int c = 0; c < 3; c++) {
        free(G[c]); free(lg[c]); free(l4[c]);
        for (int s = 0; s < 3; s++) free(lv[s][c]);
    }
}

/* Loss = energy of the final field in the shell r_lo <= K < r_hi (energy
 * weights 1/2 on kz = 0), after nsteps RK4 steps from V0; grad = dLoss/dV0
 * (complex: dL/dRe + i dL/dIm of the stored coefficients).  Forcing and the
 * reality condition are off.  Stores every state (CPU reference; sizes are
 * small). */
double or_adjoint_shell_energy(const or_grid *g0, const or_fluid *f, double *const V0[3], int nsteps, double r_lo,
                               double r_hi, double *grad[3])
{
    or_grid gg = *g0;
    gg.strong_reality = 0;
    const or_grid *g = &gg;
    const size_t nc = ncplx(g);
    const int nzc = g->nz / 2 + 1;
    double **hist = (double **)malloc(sizeof(double *) * 3 * (size_t)(nsteps + 1));
    for (int c = 0; c < 3; c++) { hist[c] = (double *)malloc(sizeof(cplx) * nc); memcpy(hist[c], V0[c], sizeof(cplx) * nc); }
    or_fluid fl = *f;
    fl.force_kind = 0;
    for (int n = 1; n <= nsteps; n++) {
        double *W[3];
        for (int c = 0; c < 3; c++) {
            W[c] = hist[3 * n + c] = (double *)malloc(sizeof(cplx) * nc);
            memcpy(W[c], hist[3 * (n - 1) + c], sizeof(cplx) * nc);
        }
        or_rk4_step(g, &fl, W);
    }
    double loss = 0.0;
    for (int c = 0; c < 3; c++) {
        const cplx *A = (const cplx *)hist[3 * nsteps + c];
        cplx *gr = (cplx *)grad[c];
        for (int lx = 0; lx < g->nx; lx++)
            for (int ly = 0; ly < g->ny; ly++)
                for (int lz = 0; lz < nzc; lz++) {
                    size_t i = cidx(g, lx, ly, lz);
                    double K = or_kmag(g, lx, ly, lz);
                    double w = lz == 0 ? 0.5 : 1.0;
                    if (K >= r_lo && K < r_hi) {
                        loss += w * (creal(A[i]) * creal(A[i]) + cimag(A[i]) * cimag(A[i]));
                        gr[i] = 2 * w * A[i];
                    } else {
                        gr[i] = 0;
                    }
                }
    }
    for (int n = nsteps - 1; n >= 0; n--) or_rk4_step_vjp(g, &fl, &hist[3 * n], grad);
    for (int q = 0; q < 3 * (nsteps + 1); q++) free(hist[q]);
    free(hist);
    return loss;
}

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                              */
/* ------------------------------------------------------------------------ */

/* Get_local_energy (FFF_pencil_energy.cc:59-67): sum|A|^2 - sum_{kz=0}|A|^2/2 */
double or_energy(const or_grid *g, const double *spec)
{
    const cplx *A = (const cplx *)spec;
    const size_t nc = ncplx(g);
    const int nzc = g->nz / 2 + 1;
    double s = 0.0, s0 = 0.0;
    for (size_t i = 0; i < nc; i++) s += creal(A[i]) * creal(A[i]) + cimag(A[i]) * cimag(A[i]);
    for (size_t i = 0; i < nc; i += nzc) s0 += creal(A[i]) * creal(A[i]) + cimag(A[i]) * cimag(A[i]);
    return s - s0 / 2;
}

/* Get_local_Sn (universal_energy.cc:106-130) */
double or_Sn(const or_grid *g, const double *spec, int n)
{
    const cplx *A = (const cplx *)spec;
    const int nzc = g->nz / 2 + 1;
    const int kmax = or_min_radius_outside(g);
    double s = 0.0;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K = or_kmag(g, lx, ly, lz);
                if (K <= kmax) {
                    cplx a = A[cidx(g, lx, ly, lz)];
                    s += or_multiplicity(g, lx, ly, lz) * pow(K, n) * (creal(a) * creal(a) + cimag(a) * cimag(a));
                }
            }
    if (n == 0) s -= or_multiplicity(g, 0, 0, 0) * (creal(A[0]) * creal(A[0]) + cimag(A[0]) * cimag(A[0]));
    return s;
}

/* Compute_local_shell_spectrum + Compute_shell_spectrum (universal_energy.cc:
 * 279-360), 3D Stepanov normalisation.  Returns the number of bins written
 * (Min_radius_outside()+1). */
int or_shell_spectrum(const or_grid *g, const double *spec, int n, double *Sk, int nbins)
{
    const cplx *A = (const cplx *)spec;
    const int nzc = g->nz / 2 + 1;
    const int kmax = or_min_radius_outside(g);
    if (nbins < kmax + 1) return -1;
    double *cnt = (double *)calloc((size_t)nbins, sizeof(double));
    for (int i = 0; i < nbins; i++) Sk[i] = 0.0;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K = or_kmag(g, lx, ly, lz);
                int idx = (int)ceil(K);
                if (idx <= kmax) {
                    double m = or_multiplicity(g, lx, ly, lz);
                    cplx a = A[cidx(g, lx, ly, lz)];
                    Sk[idx] += m * pow(K, n) * (creal(a) * creal(a) + cimag(a) * cimag(a)) * (K * K);
                    cnt[idx] += 2 * m;
                }
            }
    const double kf = g->kfactor[0] * g->kfactor[1] * g->kfactor[2];
    for (int i = 0; i <= kmax; i++) Sk[i] = cnt[i] <= 0 ? 0.0 : Sk[i] * (4 * M_PI) / (cnt[i] * kf);
    free(cnt);
    return kmax + 1;
}

/* Fill_array_shell (universal_ET.cc:880-892) */
void or_fill_array_shell(const or_grid *g, const double *spec, double *out, double r_in, double r_out)
{
    const cplx *A = (const cplx *)spec;
    cplx *B = (cplx *)out;
    const int nzc = g->nz / 2 + 1;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K = or_kmag(g, lx, ly, lz);
                size_t i = cidx(g, lx, ly, lz);
                B[i] = (K > r_in && K <= r_out) ? A[i] : 0;
            }
}

/* Local_shell_mult_single (universal_ET.cc:51-71) */
double or_shell_mult_single(const or_grid *g, const double *a, const double *b, double r_in, double r_out)
{
    const cplx *A = (const cplx *)a, *B = (const cplx *)b;
    const int nzc = g->nz / 2 + 1;
    double s = 0.0;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K = or_kmag(g, lx, ly, lz);
                if (K > r_in && K <= r_out) {
                    size_t i = cidx(g, lx, ly, lz);
                    s += 2 * or_multiplicity(g, lx, ly, lz) * creal(A[i] * conj(B[i]));
                }
            }
    return s;
}

/* Get_shell_index (shell_etc_indices.cc:49-52): first i with radii[i] >= K */
static int shell_index(double K, const double *radii, int nradii)
{
    for (int i = 0; i < nradii; i++)
        if (radii[i] >= K) return i;
    return nradii - 1;
}

/* Local_shell_mult_all, vector form (universal_ET.cc:145-173); result has
 * nradii entries, receivers with K > Max_radius_inside dropped (:159-166). */
void or_shell_mult_all(const or_grid *g, double *const A[3], double *const B[3], const double *radii,
                       int nradii, double *result)
{
    const int nzc = g->nz / 2 + 1;
    const int kin = or_max_radius_inside(g);
    for (int i = 0; i < nradii; i++) result[i] = 0.0;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K = or_kmag(g, lx, ly, lz);
                if (K <= kin) {
                    int s = shell_index(K, radii, nradii);
                    size_t i = cidx(g, lx, ly, lz);
                    double d = 0.0;
                    for (int c = 0; c < 3; c++) {
                        cplx x = ((const cplx *)A[c])[i], y = ((const cplx *)B[c])[i];
                        d += creal(x) * creal(y) + cimag(x) * cimag(y);
                    }
                    result[s] += 2 * or_multiplicity(g, lx, ly, lz) * d;
                }
            }
}

/* Default log-spaced radii (Global_process_advanced_vars.cc:137-175): n+1
 * entries, radii[n] = infinity. */
void or_default_radii(const or_grid *g, int n, double *radii)
{
    int mri = or_max_radius_inside(g);
    for (int i = 0; i <= n; i++) radii[i] = 0.0;
    radii[0] = 0.0;
    if (n >= 1) radii[1] = 2.0;
    if (n >= 2) radii[2] = 4.0;
    if (n >= 3) radii[3] = 8.0;
    if (n >= 2) radii[n - 2] = mri / 2.0;
    if (n >= 1) radii[n - 1] = mri;
    radii[n] = INFINITY;
    if (n > 6) {
        double s = log2(mri / 16.0) / (n - 5);
        for (int i = 4; i <= n - 3; i++) radii[i] = 8 * pow(2, s * (i - 3));
    }
}

/* EnergyTr::Compute_flux(FluidVF&) (compute_flux.cc:51-65) with Fill_in_sphere
 * (fill_space.cc:56-64) and Prod_out_sphere_nlinV (prod_nlinV.cc:61-69).
 * flux has nspheres+1 entries; flux[0] = 0.  radii has nspheres+1 entries. */
void or_flux(const or_grid *g, double *V[3], const double *radii, int nspheres, double *flux)
{
    const size_t nc = ncplx(g);
    double *W[3], *N[3];
    for (int c = 0; c < 3; c++) {
        W[c] = (double *)malloc(sizeof(cplx) * nc);
        N[c] = (double *)malloc(sizeof(cplx) * nc);
    }
    flux[0] = 0.0;
    for (int s = 1; s <= nspheres; s++) {
        for (int c = 0; c < 3; c++) or_fill_array_shell(g, V[c], W[c], 0.0, radii[s]);
        or_compute_nlin_giver(g, V, W, N);
        double p = 0.0;
        for (int c = 0; c < 3; c++) p += or_shell_mult_single(g, N[c], V[c], radii[s], INFINITY);
        flux[s] = -p;
    }
    for (int c = 0; c < 3; c++) { free(W[c]); free(N[c]); }
}

/* EnergyTr::Compute_shell_tr(FluidVF&) (compute_shell_tr.cc:57-75) with
 * Fill_shell (fill_space.cc:150).  S is (nshells+1)x(nshells+1) row-major,
 * row = giver shell m (1..nshells-1), column = receiver shell n. */
void or_shell_to_shell(const or_grid *g, double *V[3], const double *radii, int nshells, double *S)
{
    const size_t nc = ncplx(g);
    const int nr = nshells + 1;
    double *W[3], *N[3];
    double *row = (double *)malloc(sizeof(double) * (size_t)nr);
    for (int c = 0; c < 3; c++) {
        W[c] = (double *)malloc(sizeof(cplx) * nc);
        N[c] = (double *)malloc(sizeof(cplx) * nc);
    }
    for (int i = 0; i < nr * nr; i++) S[i] = 0.0;
    for (int m = 1; m < nshells; m++) {
        for (int c = 0; c < 3; c++) or_fill_array_shell(g, V[c], W[c], radii[m - 1], radii[m]);
        or_compute_nlin_giver(g, V, W, N);
        or_shell_mult_all(g, N, V, radii, nr, row);
        for (int n = 0; n < nr; n++) S[(size_t)m * nr + n] = -row[n];
    }
    for (int c = 0; c < 3; c++) { free(W[c]); free(N[c]); }
    free(row);
}

/* ------------------------------------------------------------------------ */
/* Seeded initial condition (this build's own definition; the reference's  */
/* generator is seeded by time(0), init_cond_energy.cc:76)                   */
/* ------------------------------------------------------------------------ */

static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

static inline double uniform01(uint64_t seed, uint64_t ctr)
{
    uint64_t h = splitmix64(splitmix64(seed) ^ ctr);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

/* spectrum_kind 0: E(K) = K^4 exp(-2 (K/k0)^2), Gaussian complex amplitudes on
 * the two Craya directions; spectrum_kind 1: E(K) = K^(-5/3) for
 * kmin <= K <= kmax, unit-modulus random phases on the two Craya directions.
 * |u(k)|^2 = E(K)/(2 pi K^2) before the total is rescaled to `energy`
 * (measured with or_energy).  K=0 and the Nyquist planes kx=Nx/2, ky=Ny/2,
 * kz=Nz/2 are zero; the kz=0 plane is made Hermitian with the strong
 * reality fill. */
void or_init_field(const or_grid *g, int spectrum_kind, double k0, double kmin, double kmax,
                   double energy, uint64_t seed, double *V[3])
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nzc = nz / 2 + 1;
    cplx *v1 = (cplx *)V[0], *v2 = (cplx *)V[1], *v3 = (cplx *)V[2];
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                size_t i = cidx(g, lx, ly, lz);
                int kx = get_kx(g, lx), ky = get_ky(g, ly);
                double Kx = kx * g->kfactor[0], Ky = ky * g->kfactor[1], Kz = lz * g->kfactor[2];
                double K = sqrt(Kx * Kx + Ky * Ky + Kz * Kz);
                double Ek = 0.0;
                if (K > 0.0) {
                    if (spectrum_kind == 0) Ek = pow(K, 4) * exp(-2.0 * (K / k0) * (K / k0));
                    else if (K >= kmin && K <= kmax) Ek = pow(K, -5.0 / 3.0);
                }
                if (kx == nx / 2 || ky == ny / 2 || lz == nz / 2) Ek = 0.0;
                if (Ek == 0.0) { v1[i] = 0; v2[i] = 0; v3[i] = 0; continue; }
                double amp = sqrt(Ek / (2.0 * M_PI * K * K));
                /* Craya basis: e1 = K x z / |K x z| (x if K || z), e2 = K x e1 / |K| */
                double kp = sqrt(Kx * Kx + Ky * Ky);
                double e1[3], e2[3];
                if (kp > 0.0) { e1[0] = Ky / kp; e1[1] = -Kx / kp; e1[2] = 0.0; }
                else { e1[0] = 1.0; e1[1] = 0.0; e1[2] = 0.0; }
                e2[0] = (Ky * e1[2] - Kz * e1[1]) / K;
                e2[1] = (Kz * e1[0] - Kx * e1[2]) / K;
                e2[2] = (Kx * e1[1] - Ky * e1[0]) / K;
                uint64_t base = (uint64_t)i * 4u;
                cplx a1, a2;
                if (spectrum_kind == 0) {
                    double u1 = 1.0 - uniform01(seed, base + 0), u2 = uniform01(seed, base + 1);
                    double u3 = 1.0 - uniform01(seed, base + 2), u4 = uniform01(seed, base + 3);
                    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
                    a1 = r1 * (cos(2 * M_PI * u2) + I * sin(2 * M_PI * u2)) / 2.0;
                    a2 = r2 * (cos(2 * M_PI * u4) + I * sin(2 * M_PI * u4)) / 2.0;
                } else {
                    double p1 = uniform01(seed, base + 0), p2 = uniform01(seed, base + 1);
                    a1 = (cos(2 * M_PI * p1) + I * sin(2 * M_PI * p1)) / sqrt(2.0);
                    a2 = (cos(2 * M_PI * p2) + I * sin(2 * M_PI * p2)) / sqrt(2.0);
                }
                v1[i] = amp * (a1 * e1[0] + a2 * e2[0]);
                v2[i] = amp * (a1 * e1[1] + a2 * e2[1]);
                v3[i] = amp * (a1 * e1[2] + a2 * e2[2]);
            }
    for (int c = 0; c < 3; c++) or_strong_reality(g, V[c]);
    double e = or_energy(g, V[0]) + or_energy(g, V[1]) + or_energy(g, V[2]);
    if (e > 0.0) {
        double s = sqrt(energy / e);
        const size_t nc = ncplx(g);
        for (int c = 0; c < 3; c++)
            for (size_t i = 0; i < nc; i++) ((cplx *)V[c])[i] *= s;
    }
}
