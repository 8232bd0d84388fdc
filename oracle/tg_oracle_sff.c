/*
 * tg_oracle_sff.c -- FP64 CPU restatement of Tarang's SFF basis (sin/cos
 * along x, Fourier along y and z) and of the Rayleigh-Benard (RBC) step on it.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker), like tg_oracle.c.
 *
 * Conventions.  The reference's sin/cos transforms live in the absent fftk
 * library, so the sample grid and scaling are this build's choice, fixed by
 * what the reference's own SFF code implies:
 *   * SFF_pencil_energy.cc:88-104 (E = 2 sum|A|^2 - sum_{kz=0} - sum_{kx=0}
 *     + sum_{kx=0,kz=0}/2, equal to sum Ar^2/(2 N) in real space, :76-85)
 *     fixes the expansion f(x) = sum_m w_m A_m phi_m(x) with w_0 = 1,
 *     w_m = 2, phi = sin(m pi x / L) ('S') or cos(m pi x / L) ('C');
 *   * SFF_pencil_tr.cc:84-105 (d/dx: 'S' -> +Kx A, 'C' -> -Kx A with
 *     Kx = lx pi / L, SFF_pencil_inline.h:63, SFF_pencil.cc:48-56) agrees.
 * Sample grid: x_j = (j + 1/2) L / Nx, j = 0..Nx-1 (cell centres, no sample
 * on the walls).  Forward: A_m = (1/Nx) sum_j f_j phi_m(x_j) (DCT-II / DST-II
 * over 2 Nx); the sine mode m = Nx has no slot and is dropped, A_0 = 0 for
 * 'S'.  Inverse: f_j = sum_m w_m A_m phi_m(x_j).  Along y, z: the FFF
 * transforms (r2c along z, complex along y), forward scaled 1/(Ny Nz).
 * Parity unpinned by any reference vector (fftk absent); pinned here by
 * scipy.fft.dct/dst type-2 golden fixtures, exact derivatives of single
 * modes, Parseval against the reference energy formula, and round trips.
 *
 * Parities (basis_table, Global.cc:138-146, sincostr_switch "SFF"):
 *   Vx 'S', Vy 'C', Vz 'C', divergence and Vi^2 'C', VxVy 'S', VxVz 'S',
 *   VyVz 'C'; the temperature takes the Vx parity ('S', Global_process_
 *   basic_vars.cc:108), VxT 'C' (Vi2), VyT 'S' (VxVy), VzT 'S' (VxVz).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tg_oracle_int.h"

typedef or_cplx cplx;

static inline size_t sidx(const or_grid *g, int lx, int ly, int lz)
{
    return ((size_t)lx * g->ny + ly) * (size_t)(g->nz / 2 + 1) + lz;
}
static inline int sky(const or_grid *g, int ly) { return ly <= g->ny / 2 ? ly : ly - g->ny; }

/* K components (SFF_pencil_inline.h:63,73,93: kx = lx, ky Fourier, kz = lz) */
static inline void sff_K(const or_grid *g, int lx, int ly, int lz, double K[3])
{
    K[0] = lx * g->kfactor[0];
    K[1] = sky(g, ly) * g->kfactor[1];
    K[2] = lz * g->kfactor[2];
}

double or_sff_kmag(const or_grid *g, int lx, int ly, int lz)
{
    double K[3];
    sff_K(g, lx, ly, lz, K);
    return sqrt(K[0] * K[0] + K[1] * K[1] + K[2] * K[2]);
}

/* phi_m(x_j) table, N x N, row m */
static double *sc_table(int n, int parity)
{
    double *t = (double *)malloc(sizeof(double) * (size_t)n * n);
    for (int m = 0; m < n; m++)
        for (int j = 0; j < n; j++) {
            /* the phase pi m (2j+1) / (2n) reduced modulo 2 pi in integers first:
             * sin and cos of a large rounded argument would carry its rounding
             * error (|a| eps, 1e-13 at n = 768) into every entry */
            const long r = ((long)m * (2 * j + 1)) % (4 * (long)n);
            const double a = M_PI * (double)r / (2.0 * n);
            t[(size_t)m * n + j] = parity == OR_SIN ? sin(a) : cos(a);
        }
    return t;
}

/* Along x, per (ly, lz) column of A: forward (dir -1) or inverse (dir +1). */
static void sc_x(const or_grid *g, int parity, cplx *A, int dir)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    double *t = sc_table(nx, parity);
#pragma omp parallel
    {
        cplx *in = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for collapse(2) schedule(static)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                for (int l = 0; l < nx; l++) in[l] = A[sidx(g, l, ly, lz)];
                for (int o = 0; o < nx; o++) {
                    cplx acc = 0;
                    if (dir < 0) {  /* o = m, sum over samples j */
                        if (!(parity == OR_SIN && o == 0))
                            for (int j = 0; j < nx; j++) acc += in[j] * t[(size_t)o * nx + j];
                        acc /= nx;
                    } else {        /* o = j, sum over modes m */
                        for (int m = (parity == OR_SIN ? 1 : 0); m < nx; m++)
                            acc += (m == 0 ? 1.0 : 2.0) * in[m] * t[(size_t)m * nx + o];
                    }
                    A[sidx(g, o, ly, lz)] = acc;
                }
            }
        free(in);
    }
    free(t);
}

/* SFF_PENCIL::Forward_transform (SFF_pencil_tr.cc:43-52) */
void or_sff_forward(const or_grid *g, int parity, const double *real_in, double *spec_out)
{
    cplx *A = (cplx *)spec_out;
    or_int_z_r2c(g, real_in, A);
    or_int_fft_y(g, A, -1);
    const double s = 1.0 / ((double)g->ny * (double)g->nz);
    const size_t tot = (size_t)g->nx * g->ny * (g->nz / 2 + 1);
    for (size_t i = 0; i < tot; i++) A[i] *= s;
    sc_x(g, parity, A, -1);
}

/* SFF_PENCIL::Inverse_transform (SFF_pencil_tr.cc:58-67), input preserved */
void or_sff_inverse(const or_grid *g, int parity, const double *spec_in, double *real_out)
{
    const size_t tot = (size_t)g->nx * g->ny * (g->nz / 2 + 1);
    cplx *A = (cplx *)malloc(sizeof(cplx) * tot);
    memcpy(A, spec_in, sizeof(cplx) * tot);
    sc_x(g, parity, A, +1);
    or_int_fft_y(g, A, +1);
    or_int_z_c2r(g, A, real_out);
    free(A);
}

/* SFF_PENCIL::Dealias (SFF_pencil_basic.cc:94-97) via Assign_sub_array
 * (:262-292): lx in [2Nx/3+1, Nx-1], ly in [Ny/3+1, 2Ny/3-1], lz in
 * [Nz/3+1, Nz/2], empty y/z ranges clamped to their first index.  The 2/3
 * mode zeroes the union lx > 2Nx/3, |ky| > Ny/3, lz > Nz/3. */
static int sff_dealias_zero(const or_grid *g, int lx, int ly, int lz)
{
    if (g->dealias == OR_DEALIAS_REF_CORNER) {
        int ylo = g->ny / 3 + 1, yhi = 2 * g->ny / 3 - 1;
        if (yhi < ylo) yhi = ylo;
        int zlo = g->nz / 3 + 1, zhi = g->nz / 2;
        if (zhi < zlo) zhi = zlo;
        return lx >= 2 * g->nx / 3 + 1 && ly >= ylo && ly <= yhi && lz >= zlo && lz <= zhi;
    }
    if (g->dealias == OR_DEALIAS_TWOTHIRDS)
        return lx > 2 * g->nx / 3 || abs(sky(g, ly)) > g->ny / 3 || lz > g->nz / 3;
    return 0;
}

void or_sff_dealias(const or_grid *g, double *spec)
{
    cplx *A = (cplx *)spec;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz <= g->nz / 2; lz++)
                if (sff_dealias_zero(g, lx, ly, lz)) A[sidx(g, lx, ly, lz)] = 0;
}

/* Satisfy_strong_reality_condition_in_Array (SFF_pencil_basic.cc:157-182):
 * A(lx, ky, 0) = conj A(lx, -ky, 0) for -Ny/2 < ky <= 0, then the kz = Nz/2
 * plane is zeroed; weak (:185-189): only the kz = Nz/2 plane. */
void or_sff_reality(const or_grid *g, double *spec, int mode)
{
    cplx *A = (cplx *)spec;
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    if (mode == 1)
        for (int lx = 0; lx < nx; lx++)
            for (int ly = 0; ly < ny; ly++) {
                const int ky = sky(g, ly);
                if (ky <= 0) {
                    const int lm = -ky;  /* index of -ky >= 0 */
                    A[sidx(g, lx, ly, 0)] = conj(A[sidx(g, lx, lm, 0)]);
                }
            }
    if (mode == 1 || mode == 2)
        for (int lx = 0; lx < nx; lx++)
            for (int ly = 0; ly < ny; ly++) A[sidx(g, lx, ly, nzc - 1)] = 0;
}

/* SFF_PENCIL::Get_local_energy(A, B) (SFF_pencil_energy.cc:117-133) */
double or_sff_energy(const or_grid *g, const double *a, const double *b)
{
    const cplx *A = (const cplx *)a, *B = (const cplx *)b;
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    double tot = 0, z0 = 0, x0 = 0, xz0 = 0;
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const size_t i = sidx(g, lx, ly, lz);
                const double d = creal(A[i]) * creal(B[i]) + cimag(A[i]) * cimag(B[i]);
                tot += d;
                if (lz == 0) z0 += d;
                if (lx == 0) x0 += d;
                if (lx == 0 && lz == 0) xz0 += d;
            }
    return 2 * tot - z0 - x0 + xz0 / 2;
}

/* x-derivative of a field of parity p (SFF_pencil_tr.cc:77-105):
 * 'S' -> +Kx A, 'C' -> -Kx A.  add: B += ..., else B = ... */
static void sff_xderiv(const or_grid *g, int parity, const cplx *A, cplx *B, int add)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    for (int lx = 0; lx < nx; lx++) {
        const double Kx = (parity == OR_SIN ? 1.0 : -1.0) * lx * g->kfactor[0];
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const size_t i = sidx(g, lx, ly, lz);
                const cplx v = Kx * A[i];
                if (add) B[i] += v; else B[i] = v;
            }
    }
}

/* y and z derivatives (SFF_pencil_tr.cc:118-151): B (+)= i K A */
static void sff_yzderiv(const or_grid *g, int d, const cplx *A, cplx *B, int add)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const double K = d == 1 ? sky(g, ly) * g->kfactor[1] : lz * g->kfactor[2];
                const size_t i = sidx(g, lx, ly, lz);
                const cplx v = (I * K) * A[i];
                if (add) B[i] += v; else B[i] = v;
            }
}

static void rmul(size_t n, const double *a, const double *b, double *c)
{
    for (size_t i = 0; i < n; i++) c[i] = a[i] * b[i];
}

/* Nlin_incompress::Compute_nlin_scalar (compute_nlin_SF.cc:75-142) on SFF:
 * dealias U and T in place (:79-82), inverse transforms, T.nlin =
 * FT[d_j(u_j T)] (Compute_nlin_VxT, compute_diag_offdiag.cc:116-136), then
 * Compute_nlin_diag / Compute_nlin_offdiag (:47-112) with the SFF parities. */
void or_compute_nlin_rbc(const or_grid *g, double *V[3], double *T, double *N[3], double *NT)
{
    const size_t nr = (size_t)g->nx * g->ny * (size_t)(g->nz + 2);
    const size_t nc = (size_t)g->nx * g->ny * (size_t)(g->nz / 2 + 1);
    const int par[3] = {OR_SIN, OR_COS, OR_COS};
    for (int c = 0; c < 3; c++) or_sff_dealias(g, V[c]);
    or_sff_dealias(g, T);
    double *u[3], *th = (double *)malloc(sizeof(double) * nr), *p = (double *)malloc(sizeof(double) * nr);
    for (int c = 0; c < 3; c++) {
        u[c] = (double *)malloc(sizeof(double) * nr);
        or_sff_inverse(g, par[c], V[c], u[c]);
    }
    or_sff_inverse(g, OR_SIN, T, th);
    cplx *X = (cplx *)malloc(sizeof(cplx) * nc);
    cplx *n1 = (cplx *)N[0], *n2 = (cplx *)N[1], *n3 = (cplx *)N[2], *nt = (cplx *)NT;
    /* Compute_nlin_VxT */
    rmul(nr, u[0], th, p); or_sff_forward(g, OR_COS, p, NT); sff_xderiv(g, OR_COS, nt, nt, 0);
    rmul(nr, u[1], th, p); or_sff_forward(g, OR_SIN, p, (double *)X); sff_yzderiv(g, 1, X, nt, 1);
    rmul(nr, u[2], th, p); or_sff_forward(g, OR_SIN, p, (double *)X); sff_yzderiv(g, 2, X, nt, 1);
    /* Compute_nlin_diag: Vi^2 'C' */
    rmul(nr, u[0], u[0], p); or_sff_forward(g, OR_COS, p, N[0]); sff_xderiv(g, OR_COS, n1, n1, 0);
    rmul(nr, u[1], u[1], p); or_sff_forward(g, OR_COS, p, N[1]); sff_yzderiv(g, 1, n2, n2, 0);
    rmul(nr, u[2], u[2], p); or_sff_forward(g, OR_COS, p, N[2]); sff_yzderiv(g, 2, n3, n3, 0);
    /* Compute_nlin_offdiag: VxVy 'S', VyVz 'C', VxVz 'S' */
    rmul(nr, u[0], u[1], p); or_sff_forward(g, OR_SIN, p, (double *)X);
    sff_yzderiv(g, 1, X, n1, 1); sff_xderiv(g, OR_SIN, X, n2, 1);
    rmul(nr, u[1], u[2], p); or_sff_forward(g, OR_COS, p, (double *)X);
    sff_yzderiv(g, 2, X, n2, 1); sff_yzderiv(g, 1, X, n3, 1);
    rmul(nr, u[2], u[0], p); or_sff_forward(g, OR_SIN, p, (double *)X);
    sff_yzderiv(g, 2, X, n1, 1); sff_xderiv(g, OR_SIN, X, n3, 1);
    free(X); free(p); free(th);
    for (int c = 0; c < 3; c++) free(u[c]);
}

/* Compute_force_TO_rhs_RBC (Tadvance_functions.cc:477-489) for PRLARGE /
 * ULARGE after the nonlinear terms: force (compute_force_RBC.cc:95-125:
 * F_u = (theta, 0, 0), F_T = u_x), nlin -= F (Add_force_RBC, :71-85),
 * The force is taken from the fields as they were before Compute_nlin
 * dealiased them (Force.Compute_force runs first, :479-481): Fu1 = theta,
 * FT = u_x.  Pressure p = div(nlin)/K^2 with the SFF divergence (Pressure.cc:56-62,
 * universal_basic.cc:343-351: +Kx on the 'S' x component; Array_divide_ksqr
 * :202-224 with p(0,0,0) = 0), nlin += grad p (Add_pressure_gradient,
 * Tadvance_functions.cc:150-166: divergence parity 'C' -> -Kx p), then the
 * signs of Compute_rhs_scalar (:218-223). */
void or_compute_rhs_rbc(const or_grid *g, const double *Fu1, const double *FT, double *N[3], double *NT)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    cplx *n1 = (cplx *)N[0], *n2 = (cplx *)N[1], *n3 = (cplx *)N[2], *nt = (cplx *)NT;
    const cplx *v1 = (const cplx *)FT, *th = (const cplx *)Fu1;
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const size_t i = sidx(g, lx, ly, lz);
                double K[3];
                sff_K(g, lx, ly, lz, K);
                cplx a = n1[i] - th[i], b = n2[i], c = n3[i];
                cplx div = K[0] * a;
                div += (I * K[1]) * b;
                div += (I * K[2]) * c;
                const double ks = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
                const cplx p = (lx == 0 && ly == 0 && lz == 0) ? 0 : div / ks;
                a += (-K[0]) * p;
                b += (I * K[1]) * p;
                c += (I * K[2]) * p;
                n1[i] = -a; n2[i] = -b; n3[i] = -c;
                nt[i] = -(nt[i] - v1[i]);
            }
}

static void sff_exp(const or_grid *g, cplx *A, double factor)
{
    const int nx = g->nx, ny = g->ny, nzc = g->nz / 2 + 1;
    for (int lx = 0; lx < nx; lx++)
        for (int ly = 0; ly < ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                double K[3];
                sff_K(g, lx, ly, lz, K);
                A[sidx(g, lx, ly, lz)] *= exp(factor * (K[0] * K[0] + K[1] * K[1] + K[2] * K[2]));
            }
}

static void rbc_rhs(const or_grid *g, double *V[3], double *T, double *N[3], double *NT)
{
    const size_t nc = (size_t)g->nx * g->ny * (size_t)(g->nz / 2 + 1);
    double *Fu1 = (double *)malloc(sizeof(cplx) * nc), *FT = (double *)malloc(sizeof(cplx) * nc);
    memcpy(Fu1, T, sizeof(cplx) * nc);
    memcpy(FT, V[0], sizeof(cplx) * nc);
    or_compute_nlin_rbc(g, V, T, N, NT);
    or_compute_rhs_rbc(g, Fu1, FT, N, NT);
    free(Fu1);
    free(FT);
}

/* Time_advance_incompress::RK4(U, T, P, F) (RK_SF.cc:97-215, PRLARGE branch)
 * with Single_time_step_scalar / Compute_RK4_parts on U (nu) and T (kappa)
 * (Tadvance_functions.cc:298-324, 372-376, 566-570, 590-597), then the
 * reality condition of Time_advance_step (Time_advance.cc:85-93):
 * p->reality 1 strong, 2 weak, 0 none.  nu = sqrt(Pr/Ra), kappa =
 * 1/sqrt(Pr Ra) (Iscalar_main.cc:66-68). */
void or_rk4_step_rbc(const or_grid *g, const or_rbc *prm, double *V[3], double *T)
{
    const size_t nc = (size_t)g->nx * g->ny * (size_t)(g->nz / 2 + 1);
    const double dt = prm->dt;
    const double nu = sqrt(prm->Pr / prm->Ra), kappa = 1.0 / sqrt(prm->Pr * prm->Ra);
    cplx *F[4], *F0[4], *acc[4], *Nn[4];
    for (int c = 0; c < 4; c++) {
        F[c] = (cplx *)(c < 3 ? V[c] : T);
        F0[c] = (cplx *)malloc(sizeof(cplx) * nc);
        acc[c] = (cplx *)calloc(nc, sizeof(cplx));
        Nn[c] = (cplx *)malloc(sizeof(cplx) * nc);
        memcpy(F0[c], F[c], sizeof(cplx) * nc);
    }
    double *Vp[3] = {(double *)F[0], (double *)F[1], (double *)F[2]};
    double *Np[3] = {(double *)Nn[0], (double *)Nn[1], (double *)Nn[2]};
    const double a_[4] = {0.5, 0.5, 1.0, 0.0}, b_[4] = {0.5, 0.0, 0.5, 0.0}, cc_[4] = {0.5, 0.5, 1.0, 0.0};
    const double rb[4] = {1.0, 0.5, 0.5, 0.0}, rf[4] = {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6};
    for (int s = 0; s < 4; s++) {
        rbc_rhs(g, Vp, (double *)F[3], Np, (double *)Nn[3]);
        if (s > 0 && s < 3)
            for (int c = 0; c < 4; c++) memcpy(F[c], F0[c], sizeof(cplx) * nc);
        for (int c = 0; c < 4; c++) {
            const double d = c < 3 ? nu : kappa;
            if (s < 3) {
                /* Single_time_step(a, b, c): the branches used by RK4 */
                const double a = a_[s], b = b_[s], cf = cc_[s];
                if (fabs(a - b) < 1e-15) {        /* a == b: add, then exp(a) */
                    for (size_t i = 0; i < nc; i++) F[c][i] += (cf * dt) * Nn[c][i];
                    sff_exp(g, F[c], -d * a * dt);
                } else if (fabs(b) < 1e-15) {     /* b == 0: exp(a) then add */
                    sff_exp(g, F[c], -d * a * dt);
                    for (size_t i = 0; i < nc; i++) F[c][i] += (cf * dt) * Nn[c][i];
                } else {                          /* exp(a-b), add, exp(b) */
                    sff_exp(g, F[c], -d * (a - b) * dt);
                    for (size_t i = 0; i < nc; i++) F[c][i] += (cf * dt) * Nn[c][i];
                    sff_exp(g, F[c], -d * b * dt);
                }
            }
            /* Compute_RK4_parts: N *= exp(-d K^2 b dt); acc += factor dt N */
            if (fabs(rb[s]) > 1e-15) sff_exp(g, Nn[c], -d * rb[s] * dt);
            for (size_t i = 0; i < nc; i++) acc[c][i] += (rf[s] * dt) * Nn[c][i];
        }
    }
    for (int c = 0; c < 4; c++) {
        const double d = c < 3 ? nu : kappa;
        memcpy(F[c], F0[c], sizeof(cplx) * nc);
        sff_exp(g, F[c], -d * 1.0 * dt);
        for (size_t i = 0; i < nc; i++) F[c][i] = F[c][i] + acc[c][i];
        if (prm->reality) or_sff_reality(g, (double *)F[c], prm->reality);
        free(F0[c]); free(acc[c]); free(Nn[c]);
    }
}

/* Correlation::Get_Nusselt_no (Correlation.cc:144-160), PRLARGE / ULARGE:
 * Nu = 1 + 2 sqrt(Ra Pr) E(V1, T) with the SFF energy of the pair. */
double or_nusselt(const or_grid *g, const or_rbc *prm, const double *V1, const double *T)
{
    return 1.0 + 2.0 * sqrt(prm->Ra * prm->Pr) * or_sff_energy(g, V1, T);
}

/* FluidVF::Get_dt with dt_option 1 (FluidVF.cc:546-572), as Compute_nlin
 * calls it (compute_nlin.cc:60-73): V dealiased in place, its real-space
 * field (Inverse_transform), umax_i = max |u_i| over the grid
 * (Get_local_max_real_space, FluidVF.cc:499-509), sum = umax_x / Dx +
 * umax_z / Dz (+ umax_y / Dy for Ny > 1), dt = min(Courant / sum, dt_fixed),
 * Dx_i = L_i / N_i with L_i = 2 pi / kfactor_i (FFF, FFF_pencil.cc:64-72)
 * or L_x = pi / kfactor_x (SFF, SFF_pencil.cc:77-86).  sff: the SFF basis
 * (u_x 'S', u_y, u_z 'C').  umax may be NULL. */
double or_get_dt(const or_grid *g, int sff, double *V[3], double courant, double dt_fixed, double *umax)
{
    const size_t nr = (size_t)g->nx * g->ny * (g->nz + 2);
    double *R = (double *)malloc(sizeof(double) * nr);
    double um[3];
    for (int c = 0; c < 3; c++) {
        if (sff) {
            or_sff_dealias(g, V[c]);
            or_sff_inverse(g, c == 0 ? OR_SIN : OR_COS, V[c], R);
        } else {
            or_dealias(g, V[c]);
            or_inverse(g, V[c], R);
        }
        double m = 0.0;
        for (size_t r = 0; r < (size_t)g->nx * g->ny; r++)
            for (int z = 0; z < g->nz; z++) {
                const double v = fabs(R[r * (size_t)(g->nz + 2) + z]);
                if (v > m || v != v) m = v;
            }
        um[c] = m;
        if (umax) umax[c] = m;
    }
    free(R);
    const double L[3] = {(sff ? M_PI : 2 * M_PI) / g->kfactor[0], 2 * M_PI / g->kfactor[1], 2 * M_PI / g->kfactor[2]};
    const double dx = L[0] / g->nx, dy = L[1] / g->ny, dz = L[2] / g->nz;
    double sum = um[0] / dx + um[2] / dz;
    if (g->ny > 1) sum += um[1] / dy;
    const double dt = courant / sum;
    return dt < dt_fixed ? dt : dt_fixed;
}
