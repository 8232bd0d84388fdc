/*
 * tg_oracle.h -- FP64 CPU restatement of Tarang's pseudo-spectral hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in tarang_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Layouts (reference FFF basis, FFF_pencil.cc:209-213):
 *   complex spectral array A(lx,ly,lz): shape (Nx, Ny, Nz/2+1), kz fastest,
 *     interleaved re/im doubles.
 *   real array Ar(rx,ry,rz): shape (Nx, Ny, Nz+2), z fastest, last two
 *     entries of each z-line are padding.
 *
 * Wavenumber maps (FFF_pencil_inline.h:63-111):
 *   kx = lx <= Nx/2 ? lx : lx-Nx ; ky likewise ; kz = lz.
 *   Kx = kx*kfactor[0] etc.
 *
 * Transform normalisation (inferred from FFF_pencil_energy.cc:51-67, the
 * reference's transform library being absent): forward r2c scaled by
 * 1/(Nx*Ny*Nz), inverse c2r unscaled.  The inverse ignores the imaginary
 * parts of the kz=0 and kz=Nz/2 coefficients of each z-line (it returns the
 * real part of the inverse DFT of the Hermitian extension).
 *
 * How each function is pinned is stated in the header comment of tg_oracle.c.
 */
#ifndef TG_ORACLE_H
#define TG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_DEALIAS_NONE = 0, OR_DEALIAS_REF_CORNER = 1, OR_DEALIAS_TWOTHIRDS = 2 };

typedef struct or_grid {
    int nx, ny, nz;
    double kfactor[3];
    int dealias;         /* OR_DEALIAS_* */
    int strong_reality;  /* apply the strong reality condition after every RK4 step */
} or_grid;

typedef struct or_fluid {
    double nu;           /* dissipation_coefficient */
    double dt;           /* fixed dt (dt_option 0) */
    int force_kind;      /* 0 = none, 1 = constant energy supply alpha-forcing */
    double force_eps;    /* energy injection rate */
    double force_r_in;   /* band (r_in, r_out] */
    double force_r_out;
    double nu_hyper;     /* hyper_dissipation_coefficient */
    int hyper_exponent;  /* q: exp((-nu - nu_hyper K^(q-2)) K^2 a dt); 0 = none */
    /* force_kind 2, random band forcing: helicity supply rate, random
     * interval tau, the seed of the phases */
    double force_helicity;
    double force_tau;
    unsigned long long force_seed;
} or_fluid;

/* ---- 1D FFT (unnormalised, in place), sign = -1 forward, +1 backward ---- */
int  or_fft1d(double *data, int n, int sign);

/* ---- 3D FFF transforms (fftk "FFF" semantics) ---- */
void or_forward(const or_grid *g, const double *real_in, double *spec_out);
void or_inverse(const or_grid *g, const double *spec_in, double *real_out);

/* ---- basis helpers ---- */
void or_dealias(const or_grid *g, double *spec);
void or_strong_reality(const or_grid *g, double *spec);
void or_weak_reality(const or_grid *g, double *spec);
double or_multiplicity(const or_grid *g, int lx, int ly, int lz);
double or_kmag(const or_grid *g, int lx, int ly, int lz);
int  or_min_radius_outside(const or_grid *g);
int  or_max_radius_inside(const or_grid *g);

/* ---- nonlinear term, pressure, time advance ---- */
void or_compute_nlin(const or_grid *g, double *V[3], double *N[3]);
void or_compute_nlin_giver(const or_grid *g, double *V[3], double *W[3], double *N[3]);
void or_compute_force(const or_grid *g, const or_fluid *f, double *const V[3], double *F[3]);
void or_compute_rhs(const or_grid *g, double *N[3], double *const F[3]);
void or_rk4_step(const or_grid *g, const or_fluid *f, double *V[3]);
void or_euler_step(const or_grid *g, const or_fluid *f, double *V[3]);

/* ---- MHD ---- */
void or_compute_nlin_mhd(const or_grid *g, double *V[3], double *W[3], double *NU[3], double *NW[3]);
void or_rk4_step_mhd(const or_grid *g, const or_fluid *fu, double eta, double *V[3], double *W[3]);
double or_cross_energy(const or_grid *g, const double *A, const double *B);
void or_flux_mhd(const or_grid *g, double *V[3], double *W[3], const double *radii, int nspheres, double *u2b,
                 double *b2b);

/* ---- SFF basis and Rayleigh-Benard convection (tg_oracle_sff.c) ---- */
enum { OR_SIN = 0, OR_COS = 1 };
typedef struct or_rbc {
    double Ra, Pr;       /* PRLARGE / ULARGE: nu = sqrt(Pr/Ra), kappa = 1/sqrt(Pr Ra) */
    double dt;
    int reality;         /* after each RK4 step: 0 none, 1 strong, 2 weak */
} or_rbc;
void or_sff_forward(const or_grid *g, int parity, const double *real_in, double *spec_out);
void or_sff_inverse(const or_grid *g, int parity, const double *spec_in, double *real_out);
void or_sff_dealias(const or_grid *g, double *spec);
double or_get_dt(const or_grid *g, int sff, double *V[3], double courant, double dt_fixed, double *umax);
void or_sff_reality(const or_grid *g, double *spec, int mode);
double or_sff_energy(const or_grid *g, const double *A, const double *B);
double or_sff_kmag(const or_grid *g, int lx, int ly, int lz);
void or_compute_nlin_rbc(const or_grid *g, double *V[3], double *T, double *N[3], double *NT);
void or_compute_rhs_rbc(const or_grid *g, const double *Fu1, const double *FT, double *N[3], double *NT);
void or_rk4_step_rbc(const or_grid *g, const or_rbc *p, double *V[3], double *T);
double or_nusselt(const or_grid *g, const or_rbc *p, const double *V1, const double *T);

/* ---- tangent / adjoint (no counterpart in the reference) ---- */
void or_rhs_jvp(const or_grid *g, double *const V[3], double *const dV[3], double *out[3]);
void or_rhs_vjp(const or_grid *g, double *const V[3], double *const Lam[3], double *out[3]);
void or_rk4_step_vjp(const or_grid *g, const or_fluid *f, double *const V[3], double *lam[3]);
double or_adjoint_shell_energy(const or_grid *g, const or_fluid *f, double *const V0[3], int nsteps, double r_lo,
                               double r_hi, double *grad[3]);

/* ---- diagnostics ---- */
double or_energy(const or_grid *g, const double *A);
double or_Sn(const or_grid *g, const double *A, int n);
int  or_shell_spectrum(const or_grid *g, const double *A, int n, double *Sk, int nbins);
void or_fill_array_shell(const or_grid *g, const double *A, double *B, double r_in, double r_out);
double or_shell_mult_single(const or_grid *g, const double *A, const double *B, double r_in, double r_out);
void or_shell_mult_all(const or_grid *g, double *const A[3], double *const B[3],
                       const double *radii, int nradii, double *result);
void or_flux(const or_grid *g, double *V[3], const double *radii, int nspheres, double *flux);
void or_shell_to_shell(const or_grid *g, double *V[3], const double *radii, int nshells, double *S);
void or_default_radii(const or_grid *g, int n, double *radii);

/* ---- anisotropic diagnostics (tg_oracle_aniso.c); dirn = anisotropy_dirn 1..3 ---- */
double or_aniso_kpll(const or_grid *g, int lx, int ly, int lz, int dirn);
double or_aniso_kperp(const or_grid *g, int lx, int ly, int lz, int dirn);
double or_polar_angle(const or_grid *g, int lx, int ly, int lz, int dirn);
double or_azimuthal_angle(const or_grid *g, int lx, int ly, int lz, int dirn);
void or_modes_in_shell(const or_grid *g, int imax, double *count);
int  or_aniso_max_krho_inside(const or_grid *g, int dirn);
void or_ring_spectrum(const or_grid *g, double *const A[], int ncomp, int n, int dirn, const double *angles,
                      int nsec, double *out);
void or_cyl_ring_spectrum(const or_grid *g, double *const V[3], int n, int dirn, const double *kpll_array, int nslabs,
                          double *S1, double *S2);
void or_fill_array_ring(const or_grid *g, const double *A, double *B, double r_in, double r_out, double a_lo,
                        double a_hi, int dirn);
void or_fill_array_cyl_ring(const or_grid *g, const double *A, double *B, double r_in, double r_out, double h_lo,
                            double h_hi, int dirn);
void or_ring_mult_all(const or_grid *g, double *const A[3], double *const B[3], int dirn, const double *radii,
                      int nshells, const double *angles, int nsec, double *out);
void or_cyl_ring_mult_all(const or_grid *g, double *const A[3], double *const B[3], int dirn, const double *radii,
                          int nshells, const double *kpll_array, int nslabs, double *out);
void or_ring_to_ring(const or_grid *g, double *V[3], int dirn, const double *radii, int nshells,
                     const double *angles, int nsec, double *out);
void or_cyl_ring_to_ring(const or_grid *g, double *V[3], int dirn, const double *radii, int nshells,
                         const double *kpll_array, int nslabs, double *out);

/* ---- Chebyshev (ChFF) path (tg_oracle_chff.c) ---- */
void or_chff_forward(const or_grid *g, const double *real_in, double *spec_out);
void or_chff_inverse(const or_grid *g, const double *spec_in, double *real_out);
void or_cheb_xderiv(const or_grid *g, const double *A, double *B);
void or_helmholtz_line(int N, const double *f, double *x, double lambda, double vp, double vm);
void or_helmholtz(const or_grid *g, const double *R, double *U, double lam_sup, const double vp[2],
                  const double vm[2]);
void or_chff_influence(const or_grid *g, double nu, double dt, double *pp, double *pm, double *vxp, double *vxm,
                       double *M);
void or_chff_correct(const or_grid *g, const double *pp, const double *pm, const double *vxp, const double *vxm,
                     const double *M, double *P, double *V1);
void or_compute_nlin_chff(const or_grid *g, double *const V[3], double *N[3]);
void or_bdf1_chff(const or_grid *g, double nu, double dt, double *V[3], double *P);

/* ---- seeded initial condition (this build's own definition) ---- */
void or_init_field(const or_grid *g, int spectrum_kind, double k0, double kmin, double kmax,
                   double energy, uint64_t seed, double *V[3]);

#ifdef __cplusplus
}
#endif
#endif
