/*
 * tg_oracle_aniso.c -- FP64 CPU restatement of the reference's anisotropic
 * diagnostics on the FFF basis: ring and cylindrical-ring spectra and the
 * ring-to-ring and cylindrical-ring-to-ring energy transfers.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker), like tg_oracle.c.
 *
 * Pinning (tests/test_oracle_aniso.py): single modes land in the bins the
 * geometry gives (angles computed independently with numpy's arctan2);
 * summing a ring spectrum over its sectors gives the shell-binned sum of
 * the same weights; summing the ring-to-ring transfer over giver and
 * receiver sectors gives the shell-to-shell transfer of tg_oracle.c (the
 * nonlinear term is linear in the giver, and the sectors tile [0, pi/2]);
 * the same for the cylindrical transfer over slabs against a brute-force
 * giver sum; antisymmetry of both transfers under the 2/3 rule.
 */
#include "tg_oracle.h"
#include "tg_oracle_int.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define OR_MYEPS 1e-15  /* myconstant.MYEPS (Global.cc:62) */

static inline int kx_of(const or_grid *g, int lx) { return lx <= g->nx / 2 ? lx : lx - g->nx; }
static inline int ky_of(const or_grid *g, int ly) { return ly <= g->ny / 2 ? ly : ly - g->ny; }
static inline size_t ix(const or_grid *g, int lx, int ly, int lz)
{
    return ((size_t)lx * g->ny + ly) * (size_t)(g->nz / 2 + 1) + lz;
}

/* AnisKpll / AnisKperp (FFF_pencil_inline.h:696-725) */
double or_aniso_kpll(const or_grid *g, int lx, int ly, int lz, int dirn)
{
    if (dirn == 1) return kx_of(g, lx) * g->kfactor[0];
    if (dirn == 2) return ky_of(g, ly) * g->kfactor[1];
    return lz * g->kfactor[2];
}

double or_aniso_kperp(const or_grid *g, int lx, int ly, int lz, int dirn)
{
    const double a = kx_of(g, lx) * g->kfactor[0], b = ky_of(g, ly) * g->kfactor[1], c = lz * g->kfactor[2];
    if (dirn == 1) return sqrt(b * b + c * c);
    if (dirn == 2) return sqrt(a * a + c * c);
    return sqrt(a * a + b * b);
}

/* Get_polar_angle (basicfn_inline.h:357-378) of AnisKvect_polar_angle
 * (FFF_pencil_inline.h:819-827): the angle to the anisotropy axis, [0, pi] */
double or_polar_angle(const or_grid *g, int lx, int ly, int lz, int dirn)
{
    const double kpll = or_aniso_kpll(g, lx, ly, lz, dirn), kperp = or_aniso_kperp(g, lx, ly, lz, dirn);
    double t;
    if (fabs(kperp) > OR_MYEPS) t = fabs(kpll) > OR_MYEPS ? atan(kperp / kpll) : M_PI / 2;
    else t = kpll >= 0 ? 0.0 : M_PI;
    return t >= 0 ? t : t + M_PI;
}

/* Get_azimuthal_angle (basicfn_inline.h:389-407) of AnisKvect_azimuthal_angle
 * (FFF_pencil_inline.h:838-845) with AnisKh1, AnisKh2 (:729-757): the angle
 * of (Kh1, Kh2) in [0, 2 pi), 0 on the axis */
double or_azimuthal_angle(const or_grid *g, int lx, int ly, int lz, int dirn)
{
    const double a = kx_of(g, lx) * g->kfactor[0], b = ky_of(g, ly) * g->kfactor[1], c = lz * g->kfactor[2];
    const double h1 = dirn == 1 ? b : (dirn == 2 ? c : a), h2 = dirn == 1 ? c : (dirn == 2 ? a : b);
    const double kperp = sqrt(h1 * h1 + h2 * h2);
    if (!(fabs(kperp) > OR_MYEPS)) return 0.0;
    const double t = acos(h1 / kperp);
    return h2 >= 0 ? t : 2 * M_PI - t;
}

/* Anis_max_Krho_radius_inside (FFF_pencil_inline.h:784-798) */
int or_aniso_max_krho_inside(const or_grid *g, int dirn)
{
    const double a = (g->nx / 2) * g->kfactor[0], b = (g->ny / 2) * g->kfactor[1], c = (g->nz / 2) * g->kfactor[2];
    double k;
    if (dirn == 1) k = fmin(b, c);
    else if (dirn == 2) k = fmin(a, c);
    else k = fmin(a, b);
    return (int)k;
}

/* first(array >= x) over arr[0..n] (Blitz first() on the radii/angle arrays) */
static int first_ge(const double *arr, int n, double x)
{
    for (int i = 0; i <= n; i++)
        if (arr[i] >= x) return i;
    return n;
}

/* Get_sector_index (shell_etc_indices.cc:58-66): angle folded into [0, pi/2] */
static int sector_index(double theta, const double *angles, int nsec)
{
    if (theta > M_PI / 2) theta = M_PI - theta;
    return fabs(theta) < OR_MYEPS ? 1 : first_ge(angles, nsec, theta);
}

/* Get_slab_index (shell_etc_indices.cc:93-98) */
static int slab_index(double kpll, const double *kpll_array, int nslabs)
{
    return fabs(kpll) < OR_MYEPS ? 1 : first_ge(kpll_array, nslabs, fabs(kpll));
}

static double powi(double k, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; i++) r *= k;
    return r;
}

/* Universal::Compute_local_ring_spectrum (universal_energy.cc:552-590 for
 * three components, :753-786 for a scalar): out[c][shell][sector],
 * (kmax+1) x (nsec+1) per component, kmax = Max_radius_inside. */
void or_ring_spectrum(const or_grid *g, double *const A[], int ncomp, int n, int dirn, const double *angles,
                      int nsec, double *out)
{
    const int kmax = or_max_radius_inside(g), nzc = g->nz / 2 + 1;
    const size_t per = (size_t)(kmax + 1) * (nsec + 1);
    memset(out, 0, sizeof(double) * per * ncomp);
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const double K = or_kmag(g, lx, ly, lz);
                const int shell = (int)ceil(K);
                if (!(K > OR_MYEPS && shell <= kmax)) continue;
                const int sec = sector_index(or_polar_angle(g, lx, ly, lz, dirn), angles, nsec);
                const double f = or_multiplicity(g, lx, ly, lz) * powi(K, n);
                for (int c = 0; c < ncomp; c++) {
                    const cplx a = ((const cplx *)A[c])[ix(g, lx, ly, lz)];
                    out[c * per + (size_t)shell * (nsec + 1) + sec] += f * (creal(a) * creal(a) + cimag(a) * cimag(a));
                }
            }
}

/* Universal::Compute_local_cylindrical_ring_spectrum (universal_energy.cc:
 * 868-906): S1 = sum m K^n |V_pll|^2, S2 = sum m K^n (|V|^2 - |V_pll|^2) in
 * bins (ceil(Kperp) <= Anis_max_Krho_radius_inside, slab of |Kpll|),
 * (kperp_max+1) x (nslabs+1). */
void or_cyl_ring_spectrum(const or_grid *g, double *const V[3], int n, int dirn, const double *kpll_array, int nslabs,
                          double *S1, double *S2)
{
    const int kmax = or_aniso_max_krho_inside(g, dirn), nzc = g->nz / 2 + 1;
    const size_t per = (size_t)(kmax + 1) * (nslabs + 1);
    memset(S1, 0, sizeof(double) * per);
    memset(S2, 0, sizeof(double) * per);
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const double kperp = or_aniso_kperp(g, lx, ly, lz, dirn);
                const int shell = (int)ceil(kperp);
                if (shell > kmax) continue;
                const int slab = slab_index(or_aniso_kpll(g, lx, ly, lz, dirn), kpll_array, nslabs);
                const double f = or_multiplicity(g, lx, ly, lz) * powi(or_kmag(g, lx, ly, lz), n);
                double v2 = 0.0, p2 = 0.0;
                for (int c = 0; c < 3; c++) {
                    const cplx a = ((const cplx *)V[c])[ix(g, lx, ly, lz)];
                    const double s = creal(a) * creal(a) + cimag(a) * cimag(a);
                    v2 += s;
                    if (c == dirn - 1) p2 = s;
                }
                S1[(size_t)shell * (nslabs + 1) + slab] += f * p2;
                S2[(size_t)shell * (nslabs + 1) + slab] += f * (v2 - p2);
            }
}

/* Universal::Fill_array_ring (universal_ET.cc:897-913): B = A where
 * r_in < K <= r_out and a_lo < theta <= a_hi (theta unfolded, in [0, pi]),
 * or theta = 0 for the first sector; zero elsewhere. */
void or_fill_array_ring(const or_grid *g, const double *A, double *B, double r_in, double r_out, double a_lo,
                        double a_hi, int dirn)
{
    const int nzc = g->nz / 2 + 1;
    const cplx *a = (const cplx *)A;
    cplx *b = (cplx *)B;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const size_t i = ix(g, lx, ly, lz);
                const double K = or_kmag(g, lx, ly, lz), th = or_polar_angle(g, lx, ly, lz, dirn);
                const int in = K > r_in && K <= r_out &&
                               ((th > a_lo && th <= a_hi) || (fabs(th) < OR_MYEPS && a_lo < OR_MYEPS));
                b[i] = in ? a[i] : 0;
            }
}

/* Universal::Fill_array_cylindrical_ring (universal_ET.cc:917-933): B = A
 * where r_in < |K| <= r_out (the full wavenumber magnitude, as the
 * reference writes it) and h_lo < Kpll <= h_hi (signed), or Kpll = 0 for
 * the first slab. */
void or_fill_array_cyl_ring(const or_grid *g, const double *A, double *B, double r_in, double r_out, double h_lo,
                            double h_hi, int dirn)
{
    const int nzc = g->nz / 2 + 1;
    const cplx *a = (const cplx *)A;
    cplx *b = (cplx *)B;
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const size_t i = ix(g, lx, ly, lz);
                const double K = or_kmag(g, lx, ly, lz), kp = or_aniso_kpll(g, lx, ly, lz, dirn);
                const int in = K > r_in && K <= r_out && ((kp > h_lo && kp <= h_hi) || (fabs(kp) < OR_MYEPS &&
                                                                                       h_lo < OR_MYEPS));
                b[i] = in ? a[i] : 0;
            }
}

/* Universal::Local_ring_mult_all (universal_ET.cc:205-230 for one
 * component pair; the three-component form at :245-272 adds the products of
 * the components): result[shell][sector] += 2 m Re(A . B*) over
 * 0 < K <= Max_radius_inside, (nshells+1) x (nsec+1); shell = first radius
 * >= K, sector of the folded polar angle (Compute_ring_index,
 * shell_etc_indices.cc:71-87). */
void or_ring_mult_all(const or_grid *g, double *const A[3], double *const B[3], int dirn, const double *radii,
                      int nshells, const double *angles, int nsec, double *out)
{
    const int kin = or_max_radius_inside(g), nzc = g->nz / 2 + 1;
    memset(out, 0, sizeof(double) * (size_t)(nshells + 1) * (nsec + 1));
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const double K = or_kmag(g, lx, ly, lz);
                if (!(K <= kin && K > OR_MYEPS)) continue;
                const int sh = first_ge(radii, nshells, K);
                const int sec = sector_index(or_polar_angle(g, lx, ly, lz, dirn), angles, nsec);
                const size_t i = ix(g, lx, ly, lz);
                double d = 0.0;
                for (int c = 0; c < 3; c++) {
                    const cplx a = ((const cplx *)A[c])[i], b = ((const cplx *)B[c])[i];
                    d += creal(a) * creal(b) + cimag(a) * cimag(b);
                }
                out[(size_t)sh * (nsec + 1) + sec] += 2 * or_multiplicity(g, lx, ly, lz) * d;
            }
}

/* Universal::Local_cyl_ring_mult_all, three components (universal_ET.cc:
 * 335-362): result[shell][slab] += 2 m Re(A . B*) over Kperp <=
 * Anis_max_Krho_radius_inside; shell = first radius >= Kperp, slab of
 * |Kpll| (Compute_cylindrical_ring_index, shell_etc_indices.cc:103-118). */
void or_cyl_ring_mult_all(const or_grid *g, double *const A[3], double *const B[3], int dirn, const double *radii,
                          int nshells, const double *kpll_array, int nslabs, double *out)
{
    const int kmax = or_aniso_max_krho_inside(g, dirn), nzc = g->nz / 2 + 1;
    memset(out, 0, sizeof(double) * (size_t)(nshells + 1) * (nslabs + 1));
    for (int lx = 0; lx < g->nx; lx++)
        for (int ly = 0; ly < g->ny; ly++)
            for (int lz = 0; lz < nzc; lz++) {
                const double kperp = or_aniso_kperp(g, lx, ly, lz, dirn);
                if (!(kperp <= kmax)) continue;
                const int sh = first_ge(radii, nshells, kperp);
                const int sl = slab_index(or_aniso_kpll(g, lx, ly, lz, dirn), kpll_array, nslabs);
                const size_t i = ix(g, lx, ly, lz);
                double d = 0.0;
                for (int c = 0; c < 3; c++) {
                    const cplx a = ((const cplx *)A[c])[i], b = ((const cplx *)B[c])[i];
                    d += creal(a) * creal(b) + cimag(a) * cimag(b);
                }
                out[(size_t)sh * (nslabs + 1) + sl] += 2 * or_multiplicity(g, lx, ly, lz) * d;
            }
}

/* EnergyTr::Compute_ring_tr (compute_ring_tr.cc:62-77): for giver rings
 * (shell 1..nshells-1, sector 1..nsec), T(giver | receiver) = -ring_mult of
 * N(u | u_giver) with u; out[((i*(nsec+1)+j)*(nshells+1)+k)*(nsec+1)+l].
 * The reference's three-component Local_ring_mult_all adds 0 (its product
 * line is commented out, universal_ET.cc:267-268); this restatement adds
 * the product the two-array form and the cylindrical form add. */
void or_ring_to_ring(const or_grid *g, double *V[3], int dirn, const double *radii, int nshells,
                     const double *angles, int nsec, double *out)
{
    const size_t nc = (size_t)g->nx * g->ny * (g->nz / 2 + 1);
    const size_t rb = (size_t)(nshells + 1) * (nsec + 1);
    memset(out, 0, sizeof(double) * rb * rb);
    double *W[3], *N[3];
    double *row = (double *)malloc(sizeof(double) * rb);
    for (int c = 0; c < 3; c++) {
        W[c] = (double *)malloc(sizeof(cplx) * nc);
        N[c] = (double *)malloc(sizeof(cplx) * nc);
    }
    for (int i = 1; i < nshells; i++)
        for (int j = 1; j <= nsec; j++) {
            for (int c = 0; c < 3; c++)
                or_fill_array_ring(g, V[c], W[c], radii[i - 1], radii[i], angles[j - 1], angles[j], dirn);
            or_compute_nlin_giver(g, V, W, N);
            or_ring_mult_all(g, N, V, dirn, radii, nshells, angles, nsec, row);
            double *o = out + ((size_t)i * (nsec + 1) + j) * rb;
            for (size_t k = 0; k < rb; k++) o[k] = -row[k];
        }
    for (int c = 0; c < 3; c++) { free(W[c]); free(N[c]); }
    free(row);
}

/* EnergyTr::Compute_cylindrical_ring_tr (compute_cylin_ring_tr.cc:51-64) */
void or_cyl_ring_to_ring(const or_grid *g, double *V[3], int dirn, const double *radii, int nshells,
                         const double *kpll_array, int nslabs, double *out)
{
    const size_t nc = (size_t)g->nx * g->ny * (g->nz / 2 + 1);
    const size_t rb = (size_t)(nshells + 1) * (nslabs + 1);
    memset(out, 0, sizeof(double) * rb * rb);
    double *W[3], *N[3];
    double *row = (double *)malloc(sizeof(double) * rb);
    for (int c = 0; c < 3; c++) {
        W[c] = (double *)malloc(sizeof(cplx) * nc);
        N[c] = (double *)malloc(sizeof(cplx) * nc);
    }
    for (int i = 1; i < nshells; i++)
        for (int j = 1; j <= nslabs; j++) {
            for (int c = 0; c < 3; c++)
                or_fill_array_cyl_ring(g, V[c], W[c], radii[i - 1], radii[i], kpll_array[j - 1], kpll_array[j],
                                       dirn);
            or_compute_nlin_giver(g, V, W, N);
            or_cyl_ring_mult_all(g, N, V, dirn, radii, nshells, kpll_array, nslabs, row);
            double *o = out + ((size_t)i * (nslabs + 1) + j) * rb;
            for (size_t k = 0; k < rb; k++) o[k] = -row[k];
        }
    for (int c = 0; c < 3; c++) { free(W[c]); free(N[c]); }
    free(row);
}
