/*
 * tarang_gpu.h -- C ABI of the MI355X (gfx950) pseudo-spectral hot path.
 *
 * The entry points replace, on one GPU, the transform library and the
 * solver-level hot path of Tarang (anandogc/tarang).  Each declaration names
 * the reference interface it replaces (paths relative to the reference
 * checkout).
 *
 * Conventions
 *   - Every pointer argument named *_dev is a device pointer (HIP global
 *     memory).  Pointers to small result arrays (energies, spectra, fluxes)
 *     are host pointers; those calls synchronise the plan's stream.
 *   - Spectral fields: interleaved re/im FP64, shape (Nx, Ny, Nz/2+1) with
 *     kz fastest (Blitz Array<Complex,3> layout, FFF_pencil.cc:209-213).
 *   - Real fields: FP64, shape (Nx, Ny, Nz+2) with z fastest, last two
 *     entries of each z-line padding (in-place r2c layout).
 *   - A vector field is an array of 3 device pointers (V1, V2, V3).
 *   - Return value: TG_OK (0) or a negative TG_ERR_* code; the message of
 *     the last error on the calling thread is tg_last_error_string().
 *     Nothing throws across this boundary.
 *   - All work is enqueued on the plan's stream (tg_set_stream); one plan
 *     per host thread; plans are independent of each other.
 */
#ifndef TARANG_GPU_H
#define TARANG_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TG_OK 0
#define TG_ERR_ARG (-1)
#define TG_ERR_SHAPE (-2)
#define TG_ERR_OOM (-3)
#define TG_ERR_HIP (-4)
#define TG_ERR_NAN (-5)    /* a time advance wrote a non-finite value into the state */
#define TG_ERR_STATE (-6)  /* call out of sequence (adjoint backward before a forward sweep) */

/* Bases: FFF (Fourier x, y, z), SFF (sin/cos along x, Fourier y, z;
 * SFF_pencil.cc) and ChFF (Chebyshev along x on the Chebyshev-Gauss points
 * x_j = cos(pi (j + 1/2) / Nx), Fourier y, z; the basis of the reference's
 * Chebyshev solvers, time_advance/chebyshev/, whose basis class is not in
 * the reference).  For SFF, kfactor[0] = pi / L_x and strong_reality selects
 * the reality condition applied after each RK4 step: 0 none, 1 strong, 2 weak
 * (apply_strong/weak_realitycond_alltime_switch, Time_advance.cc:85-93).
 * ChFF plans are FP64 with strong_reality 0; kfactor[0] = 2 / L_x scales
 * d/dx. */
enum { TG_BASIS_FFF = 0, TG_BASIS_SFF = 1, TG_BASIS_CHFF = 2 };
/* sincos_code of tg_forward / tg_inverse (global.program.sincostr_switch[0]):
 * TG_SC_FOURIER on FFF plans; TG_SC_SIN ('S') or TG_SC_COS ('C') on SFF plans;
 * TG_SC_CHEB on ChFF plans (f(x_j) = sum_k a_k T_k(x_j)). */
enum { TG_SC_FOURIER = 0, TG_SC_SIN = 1, TG_SC_COS = 2, TG_SC_CHEB = 3 };
enum { TG_PREC_FP64 = 0, TG_PREC_FP32 = 1 };
/* Dealias modes: reference corner box (FFF_pencil_basic.cc:87-90) or the
 * two-thirds rule |kx|<=Nx/3, |ky|<=Ny/3, kz<=Nz/3. */
enum { TG_DEALIAS_NONE = 0, TG_DEALIAS_REF_CORNER = 1, TG_DEALIAS_TWOTHIRDS = 2 };
enum { TG_STEP_CONTINUE = 1 };

typedef struct tg_config {
    int basis;          /* TG_BASIS_FFF or TG_BASIS_SFF */
    int nx, ny, nz;     /* grid: nx, ny in [8,1024], nz/2 in [8,512], each 2^a, or in FP64
                           also 3*2^a, 5*2^a, 15*2^a (multiples of 8: 24, 40, 48, ..., 960);
                           other grids return TG_ERR_SHAPE */
    double kfactor[3];  /* 2 pi / L per axis (global.field.kfactor[1..3]) */
    int precision;      /* TG_PREC_FP64 or TG_PREC_FP32 (transform/product precision) */
    int dealias;        /* TG_DEALIAS_* (global.program.alias_option) */
    int strong_reality; /* apply_strong_realitycond_alltime_switch */
    int device;         /* HIP device ordinal */
} tg_config;

/* Fluid parameters (FluidVF dissipation_coefficient, fixed dt, forcing). */
typedef struct tg_fluid {
    double nu;
    double dt;
    int force_kind;     /* 0 none, 1 constant energy supply alpha-forcing, 2 random band forcing */
    double force_eps;   /* energy injection rate */
    double force_r_in;  /* forcing band (r_in, r_out] */
    double force_r_out;
    /* FluidVF hyper_dissipation_coefficient / _exponent (FluidVF.cc:72-84):
     * the viscous factors become exp((-nu - nu_hyper K^(q-2)) K^2 a dt)
     * (Universal::Array_exp_ksqr, universal_basic.cc:265-291), K^(q-2) =
     * (K^2)^((q-2)/2) with integer division; q = hyper_exponent, 0 = off,
     * else 2..16.  Taken by tg_rk4_step and the adjoint; the MHD and BDF1
     * steps return TG_ERR_ARG for q != 0. */
    double nu_hyper;
    int hyper_exponent;
    /* force_kind 2, random band forcing (Compute_force_using_random_noise_basic,
     * compute_force_random.cc:65-130): on every mode of (r_in, r_out],
     * helical amplitudes sqrt((e_k +- h_k/K) / force_tau) with phases drawn
     * from force_seed (own counter-based generator), through the Craya basis
     * to Cartesian components about z (force_amp_phase.cc:119-170,
     * FFF_pencil_inline.h:234-248); e_k = force_eps / n / modes_in_shell(i),
     * h_k likewise from force_helicity, i = ceil(K), n = the number of shells
     * floor(r_in)+1 .. ceil(r_out) (universal_basic.cc:104-143).  Needs
     * force_tau > 0 and |force_helicity| <= r_in force_eps.  The force is a
     * function of force_seed: the caller changes the seed every random
     * interval (the reference reseeds when the time passes random_next,
     * compute_force_random.cc:143-167). */
    double force_helicity;
    double force_tau;
    unsigned long long force_seed;
} tg_fluid;

/* Rayleigh-Benard parameters, Pr_option PRLARGE with Uscaling ULARGE
 * (Iscalar_main.cc:66-68: nu = sqrt(Pr/Ra), kappa = 1/sqrt(Pr Ra);
 * compute_force_RBC.cc:95-125: F_u = (theta, 0, 0), F_theta = u_x). */
typedef struct tg_rbc {
    double Ra;
    double Pr;
    double dt;
} tg_rbc;

typedef struct tg_plan tg_plan;

const char *tg_last_error_string(void);
const char *tg_version(void);

/* Replaces fftk.Init("FFF", Nx, Ny, Nz, num_p_rows, comm) (FFF_pencil.cc:91)
 * and the FFF_PENCIL constructor's shape queries (:93-94). */
int tg_plan_create(const tg_config *cfg, tg_plan **plan);
int tg_plan_destroy(tg_plan *plan);
int tg_set_stream(tg_plan *plan, void *hip_stream);
int tg_synchronize(tg_plan *plan);
/* complex_shape = (Nx, Ny, Nz/2+1), real_shape = (Nx, Ny, Nz+2)
 * (fftk.Get_FA_shape / Get_RA_shape, FFF_pencil.cc:93-94). */
int tg_get_shapes(const tg_plan *plan, int64_t complex_shape[3], int64_t real_shape[3]);
/* Bytes of device memory held by the plan. */
int64_t tg_plan_device_bytes(const tg_plan *plan);

/* FFF_PENCIL::Forward_transform -> fftk.Forward_transform_3d("FFF", Ar, A)
 * (FFF_pencil_tr.cc:43-51): r2c, scaled by 1/(Nx Ny Nz); sincos_code
 * TG_SC_FOURIER.  SFF_PENCIL::Forward_transform (SFF_pencil_tr.cc:43-52):
 * sin ('S') or cos ('C') transform along x on the cell-centred samples
 * x_j = (j + 1/2) L_x / Nx, A_m = (1/Nx) sum_j f_j phi_m(x_j) (the sine mode
 * m = Nx dropped, A_0 = 0 for 'S'), r2c along z and FFT along y scaled
 * 1/(Ny Nz). */
int tg_forward(tg_plan *plan, int sincos_code, const double *real_dev, double *spec_dev);
/* FFF_PENCIL::Inverse_transform (FFF_pencil_tr.cc:57-66): c2r, unscaled,
 * input preserved.  SFF (SFF_pencil_tr.cc:58-67): f_j = sum_m w_m A_m
 * phi_m(x_j), w_0 = 1, w_m = 2. */
int tg_inverse(tg_plan *plan, int sincos_code, const double *spec_dev, double *real_dev);

/* FluidVF::Get_dt with dt_option 1 (FluidVF.cc:546-572), as Compute_nlin
 * calls it at the start of a step (compute_nlin.cc:60-73): V dealiased in
 * place, umax_i = max |u_i| over the real-space grid (Inverse_transform +
 * Get_local_max_real_space, FluidVF.cc:499-509), dt = min(courant / (umax_x
 * / Dx + umax_y / Dy + umax_z / Dz), dt_fixed), D_i = L_i / N_i with L_i =
 * 2 pi / kfactor_i (FFF_pencil.cc:64-72; SFF: L_x = pi / kfactor_x,
 * SFF_pencil.cc:77-86, u_x a sine and u_y, u_z cosine series).  FFF and
 * SFF plans; umax (3 doubles) may be null.  Synchronises the plan's stream;
 * TG_ERR_NAN if a maximum is not finite.  dt_option 0 is dt_fixed itself. */
int tg_get_dt(tg_plan *plan, double *const V_dev[3], double courant, double dt_fixed, double *dt, double umax[3]);

/* FFF_PENCIL::Dealias (FFF_pencil_basic.cc:87-90) in place, per the plan's mode. */
int tg_dealias(tg_plan *plan, double *const V_dev[], int ncomp);
/* Satisfy_strong_reality_condition_in_Array (FFF_pencil_basic.cc:139-179;
 * SFF: SFF_pencil_basic.cc:157-182). */
int tg_strong_reality(tg_plan *plan, double *const V_dev[], int ncomp);
/* Satisfy_weak_reality_condition_in_Array (FFF_pencil_basic.cc:181-186,
 * SFF_pencil_basic.cc:185-189): the kz = Nz/2 plane is zeroed. */
int tg_weak_reality(tg_plan *plan, double *const V_dev[], int ncomp);

/* Nlin_incompress::Compute_nlin(FluidVF&) (compute_nlin.cc:58-109):
 * dealiases V in place, nlin_i = FT[d_j(u_j u_i)]. */
int tg_compute_nlin(tg_plan *plan, double *const V_dev[3], double *const nlin_dev[3]);
/* Time_advance_incompress::Compute_force_TO_rhs (Tadvance_functions.cc:441-453):
 * nlin_i = -(N_i - F_i + i K_i p), p = i K.(N - F)/K^2. */
int tg_compute_rhs(tg_plan *plan, double *const V_dev[3], const tg_fluid *fluid, double *const rhs_dev[3]);
/* Nlin_incompress::Compute_nlin(FluidVF&, PlainFluidVF&) (compute_nlin_ET_VF.cc:71-87):
 * dealiases V and W in place, nlin_i = FT[d_j(u_j w_i)]. */
int tg_compute_nlin_giver(tg_plan *plan, double *const V_dev[3], double *const W_dev[3],
                          double *const nlin_dev[3]);

/* Time_advance_incompress::Time_advance_step with integration_scheme RK4
 * (Time_advance.cc:43-65, RK.cc:91-170): nsteps Lawson RK4 steps of V in
 * place.  flags & TG_STEP_CONTINUE reuses the stage-1 transform prepared by
 * the previous call on the same V and the same forcing band (valid unless V
 * was modified outside this library since).  Returns TG_ERR_NAN when the
 * state holds a non-finite value after the call (the reference stops on a
 * NaN energy, Ifluid_main.cc:127-128); the call waits for its work to end. */
int tg_rk4_step(tg_plan *plan, double *const V_dev[3], const tg_fluid *fluid, int nsteps, int flags);

/* MHD, Elsasser form.  Nlin_incompress::Compute_nlin(FluidVF& U, FluidVF& W)
 * (compute_nlin_VF_MHD.cc:71-123): dealiases V (u) and W (b) in place;
 * nlinV = (N+ + N-)/2 = FT[(u.grad)u - (b.grad)b], nlinW = (N+ - N-)/2
 * = FT[(u.grad)b - (b.grad)u] in divergence form. */
int tg_compute_nlin_mhd(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], double *const nlinV_dev[3],
                        double *const nlinW_dev[3]);
/* Time_advance_incompress::RK4(U, W, P, F) (RK_VF_MHD.cc:126-190), nu =
 * fluid->nu, eta for the magnetic field; no forcing (fluid->force_kind 0).
 * Returns TG_ERR_NAN when u or b holds a non-finite value after the call. */
int tg_rk4_step_mhd(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], const tg_fluid *fluid, double eta,
                    int nsteps);
/* Correlation::Get_cross_helicity (Correlation.cc:225-230). */
int tg_cross_helicity(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], double *hc);
/* EnergyTr::Compute_flux(FluidVF& U, FluidVF& W) (compute_flux.cc:158-211):
 * u2b = flux_VF_Uin_Wout, b2b = flux_VF_Win_Wout.  Dealiases V and W. */
int tg_flux_mhd(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], const double *radii, int nspheres,
                double *u2b, double *b2b);

/* Rayleigh-Benard convection on an SFF plan (state u_x 'S', u_y, u_z 'C',
 * theta 'S'; parities from basis_table, Global.cc:138-146).
 * Nlin_incompress::Compute_nlin(FluidVF& U, FluidSF& T) ->
 * Compute_nlin_scalar (compute_nlin_SF.cc:75-142): dealiases V and T in
 * place; nlin_i = FT[d_j(u_j u_i)], nlinT = FT[d_j(u_j theta)]. */
int tg_compute_nlin_rbc(tg_plan *plan, double *const V_dev[3], double *T_dev, double *const nlin_dev[3],
                        double *nlinT_dev);
/* Compute_force_TO_rhs_RBC (Tadvance_functions.cc:477-489): the force from
 * V and T as given, then the nonlinear terms (V, T dealiased in place),
 * pressure and signs: rhs = -(N - F + grad p), rhsT = -(N_T - u_x). */
int tg_compute_rhs_rbc(tg_plan *plan, double *const V_dev[3], double *T_dev, const tg_rbc *rbc,
                       double *const rhs_dev[3], double *rhsT_dev);
/* Time_advance_incompress::RK4(U, T, P, F) (RK_SF.cc:97-215) followed by the
 * plan's reality condition: nsteps Lawson RK4 steps of (V, T) in place.
 * Returns TG_ERR_NAN when the state holds a non-finite value after the call. */
int tg_rk4_step_rbc(tg_plan *plan, double *const V_dev[3], double *T_dev, const tg_rbc *rbc, int nsteps);
/* Correlation::Get_Nusselt_no (Correlation.cc:144-160), PRLARGE / ULARGE:
 * Nu = 1 + 2 sqrt(Ra Pr) E(V1, T) with the SFF pair energy. */
int tg_nusselt(tg_plan *plan, const double *V1_dev, const double *T_dev, const tg_rbc *rbc, double *nu);

/* Universal::Get_total_energy (universal_energy.cc:63-71) of one component:
 * FFF sum |A|^2 - sum_{kz=0} |A|^2 / 2; SFF (SFF_pencil_energy.cc:88-104)
 * 2 sum |A|^2 - sum_{kz=0} - sum_{kx=0} + sum_{kx=0,kz=0} / 2. */
int tg_energy(tg_plan *plan, const double *A_dev, double *energy);
/* Universal::Get_total_Sn (universal_energy.cc:106-141): sum m K^n |A|^2. */
int tg_Sn(tg_plan *plan, const double *A_dev, int n, double *Sn);
/* Universal::Compute_shell_spectrum (universal_energy.cc:279-360), Stepanov
 * normalisation; nbins >= Min_radius_outside()+1. */
int tg_shell_spectrum(tg_plan *plan, const double *A_dev, int n, double *Sk, int nbins);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tg_shell_spectrum without a plan and without any
 * allocation.  The grid, kfactor and device come from cfg (FFF basis; other
 * fields ignored); A_dev is the complex field (Nx, Ny, Nz/2+1); Sk_dev
 * (device memory, nbins >= Min_radius_outside + 1 doubles) receives the same
 * E(k) as tg_shell_spectrum, bit for bit.  The block partial sums live in a
 * static 16 MiB device buffer of the library, one per GPU, held until the
 * process ends: grids whose partials exceed it (about 590^3 at kfactor 1)
 * return TG_ERR_ARG.  Asynchronous on `stream` (a hipStream_t, null for the
 * default stream); makes cfg->device current.  Calls that overlap on one GPU
 * share the buffer and race: serialise them (one stream, or events). */
int tg_shell_spectrum_noalloc(const tg_config *cfg, const double *A_dev, int n, double *Sk_dev, int nbins,
                              void *stream);
/* EnergyTr::Compute_flux(FluidVF&) (compute_flux.cc:51-65): flux[0..nspheres],
 * radii[0..nspheres].  Dealiases V in place. */
int tg_flux(tg_plan *plan, double *const V_dev[3], const double *radii, int nspheres, double *flux);
/* EnergyTr::Compute_shell_tr(FluidVF&) (compute_shell_tr.cc:57-75):
 * S[(nshells+1)*(nshells+1)], row = giver shell.  Dealiases V in place. */
int tg_shell_to_shell(tg_plan *plan, double *const V_dev[3], const double *radii, int nshells, double *S);
/* Both at once when the flux spheres are the shell radii (no single
 * counterpart: EnergyTr::Compute_flux and Compute_shell_tr called with the
 * same radii, compute_flux.cc:51-65 and compute_shell_tr.cc:57-75): one giver
 * evaluation per shell gives the shell-to-shell row S (as tg_shell_to_shell)
 * and, summed over the givers inside each sphere and the receivers outside
 * it, the flux (as tg_flux, to round-off; the nonlinear term is linear in the
 * giver).  radii[0..n], flux[0..n], S[(n+1)*(n+1)].  Needs radii[0] = 0,
 * non-decreasing radii and radii[n] >= Min_radius_outside (a last
 * sphere enclosing every mode, e.g. inf), TG_ERR_ARG otherwise.  Dealiases V
 * in place. */
int tg_energy_transfer(tg_plan *plan, double *const V_dev[3], const double *radii, int n, double *flux, double *S);

/* ---- ChFF (Chebyshev x) path; arrays (Nx, Ny, Nz/2+1), Chebyshev
 * coefficients along x ---- */
/* d/dx in Chebyshev space: c_k b_k = b_{k+2} + 2 (k+1) a_{k+1} (the
 * derivative at the walls, influence_matrix.cc:215-229, is its sum), times
 * kfactor[0].  A and B distinct. */
int tg_cheb_xderiv(tg_plan *plan, const double *A_dev, double *B_dev);
/* Time_advance_incompress::Helmholtz_complex_full_array (helmholtz.cc:94-118)
 * with Helmholtz_complex / Tridiagonal_solver (:18-64, :155-231): per
 * (ky, kz) line, u'' - lambda u = -R (quasi-tridiagonal, Thual's algorithm,
 * even and odd coefficients apart), lambda = Kperp^2 + lambda_supplement,
 * u(+1) = value_plus1, u(-1) = value_minus1 ({re, im}).  R and U distinct. */
int tg_helmholtz(tg_plan *plan, const double *R_dev, double *U_dev, double lambda_supplement,
                 const double value_plus1[2], const double value_minus1[2]);
/* Nlin_incompress::Compute_nlin(U) on ChFF: N_i = sum_j d_j (u_j u_i), the
 * products at the Gauss points, x-derivatives by the recurrence (no
 * dealiasing). */
int tg_compute_nlin_chff(tg_plan *plan, double *const V_dev[3], double *const nlin_dev[3]);
/* Time_advance_incompress::BDF1 (BDF1.cc:63-121), unforced, nsteps steps of V
 * in place, P the pressure: p from the divergence of N with zero wall values,
 * each component from (D^2 - Kperp^2 - 1/(dt nu)) u = (N_i - u_i/dt +
 * d_i p)/(-nu) with no-slip walls, the influence-matrix correction of p and
 * u_x after u_x (Compute_pressure_Ux_chebyshev, influence_matrix.cc:137-180;
 * homogeneous solutions :74-133, computed once per (nu, dt)). */
int tg_bdf1_step(tg_plan *plan, double *const V_dev[3], double *P_dev, const tg_fluid *fluid, int nsteps);

/* Anisotropic diagnostics (FFF plans).  dirn is the anisotropy axis
 * (global.field.anisotropy_dirn, 1..3); theta is the angle of K to it, folded
 * into [0, pi/2] for binning; edge arrays hold nsub+1 non-decreasing values
 * (sector angles up to pi/2, or |Kpll| edges); bins are shell-major with
 * nsub+1 sub-bins per shell (sub-bin 0 stays empty, as in the reference).
 *
 * Universal::Compute_ring_spectrum (universal_energy.cc:552-628, scalar form
 * :753-786; Correlation::Compute_ring_spectrum, Correlation.cc:301-311):
 * out[c][shell][sector] = sum m K^n |A_c|^2 over 0 < K, shell = ceil(K) <=
 * Max_radius_inside; ncomp 1 or 3; out has ncomp (Kmax+1)(nsectors+1). */
int tg_ring_spectrum(tg_plan *plan, const double *const A_dev[], int ncomp, int n, int dirn,
                     const double *sector_angles, int nsectors, double *out);
/* Universal::Compute_cylindrical_ring_spectrum (universal_energy.cc:868-
 * 925): S1 = sum m K^n |V_pll|^2, S2 = sum m K^n (|V|^2 - |V_pll|^2) in bins
 * shell = ceil(Kperp) <= Anis_max_Krho_radius_inside, slab of |Kpll|
 * (Get_slab_index); S1, S2 each (Kperp_max+1)(nslabs+1). */
int tg_cylindrical_ring_spectrum(tg_plan *plan, const double *const V_dev[3], int n, int dirn, const double *kpll,
                                 int nslabs, double *S1, double *S2);
/* EnergyTr::Compute_ring_tr (compute_ring_tr.cc:62-77): for the giver rings
 * (shell 1..nshells-1 of radii[0..nshells], sector 1..nsectors;
 * Fill_array_ring, universal_ET.cc:897-913), the transfer into every
 * receiver ring, -sum 2m Re(N(u|u_giver) . u*) over 0 < K <=
 * Max_radius_inside (Ring_mult_all, universal_ET.cc:205-272).
 * out[((i (ns+1) + j) (nr+1) + k) (ns+1) + l], nr = nshells, ns = nsectors.
 * The reference's three-component Local_ring_mult_all adds 0 (its product
 * is commented out, universal_ET.cc:267-268); this adds the product its
 * one-component and cylindrical forms add.  Dealiases V in place. */
int tg_ring_to_ring(tg_plan *plan, double *const V_dev[3], int dirn, const double *radii, int nshells,
                    const double *sector_angles, int nsectors, double *out);
/* EnergyTr::Compute_cylindrical_ring_tr (compute_cylin_ring_tr.cc:51-64):
 * giver rings radii[i-1] < |K| <= radii[i], kpll[j-1] < Kpll <= kpll[j]
 * (Fill_array_cylindrical_ring, universal_ET.cc:917-933), receivers binned
 * by the Kperp shell and the slab of |Kpll| over Kperp <=
 * Anis_max_Krho_radius_inside (Cyl_ring_mult_all, universal_ET.cc:
 * 335-375).  Layout as tg_ring_to_ring.  Dealiases V in place. */
int tg_cylindrical_ring_to_ring(tg_plan *plan, double *const V_dev[3], int dirn, const double *radii, int nshells,
                                const double *kpll, int nslabs, double *out);

/* Per-kernel timing with HIP events on the plan's stream (no counterpart
 * in the reference, whose only timing is the total wall time printed by
 * main.cc:146-157).  Kernel classes: */
enum {
    TG_KERNEL_X_INV = 0,        /* x-inverse of the state (step start) */
    TG_KERNEL_Y_FFT = 1,        /* y passes (inverse and forward) */
    TG_KERNEL_Z_PRODUCTS = 2,   /* z c2r + real-space products + z r2c */
    TG_KERNEL_X_NLIN_RK = 3,    /* x forward + nonlinear term + RK4 stage + x inverse */
    TG_KERNEL_X_NLIN_GIVER = 4, /* x forward of the giver products */
    TG_KERNEL_REALITY = 5,      /* strong reality condition + patch */
    TG_KERNEL_Z_R2C = 6,
    TG_KERNEL_Z_C2R = 7,
    TG_KERNEL_X_FWD = 8,
    TG_KERNEL_REDUCE = 9,
    TG_KERNEL_X_INV_SYM = 10,   /* adjoint: x-inverse of the symmetrised cotangent */
    TG_KERNEL_Z_ADJ = 11,       /* adjoint: z c2r + products with the velocity tape + r2c */
    TG_KERNEL_X_FWD_ADJ = 12,   /* adjoint: x forward + RK4 stage transpose */
    TG_KERNEL_X_SC = 13,        /* SFF sin/cos x passes (transforms, stage-1 input, patches) */
    TG_KERNEL_Z_RBC = 14,       /* RBC z pass: 4 c2r, 9 products (8 in an RK stage), r2c of each */
    TG_KERNEL_X_RBC = 15,       /* RBC x pass: sin/cos forward + N, N_T + force + projection + RK4 stage */
    TG_KERNEL_COUNT = 16
};
int tg_profile_enable(tg_plan *plan, int on);
int tg_profile_reset(tg_plan *plan);
/* Synchronises the stream; total milliseconds and launch count of a class. */
int tg_profile_read(tg_plan *plan, int kernel, double *total_ms, int64_t *launches);
const char *tg_kernel_name(int kernel);

/* Adjoint (no counterpart in the reference, which has no adjoint).  The
 * inner product is the real dot product of the stored half-spectrum
 * coefficients (Re, Im), so a gradient dL/dV has components dL/dRe + i dL/dIm.
 * Supported: no forcing (fluid->force_kind 0) and strong_reality 0 (the
 * forward map is then the plain Lawson RK4 of tg_rk4_step).
 *
 * tg_rhs_vjp: out = f'(V)^T lam for f(V) = -P NL(D V), the unforced
 * right-hand side of tg_compute_rhs.  V and lam are not modified. */
int tg_rhs_vjp(tg_plan *plan, double *const V_dev[3], double *const lam_dev[3], double *const out_dev[3]);
/* Forward sweep with a tape: nsteps RK4 steps of V in place (as tg_rk4_step
 * with flags 0), keeping a checkpoint of V every checkpoint_every steps and
 * the real-space stage velocities of the last segment.  Device memory held
 * until tg_adjoint_free or plan destruction: ceil(nsteps/every) checkpoints
 * of 3 FP64 spectral components plus 12*every real fields in the transform
 * precision. */
int tg_adjoint_forward(tg_plan *plan, double *const V_dev[3], const tg_fluid *fluid, int nsteps,
                       int checkpoint_every);
/* Backward sweep over the steps of the last tg_adjoint_forward: lam holds
 * dL/dV_n on entry and dL/dV_0 on return.  Earlier segments are recomputed
 * from their checkpoints. */
int tg_adjoint_backward(tg_plan *plan, double *const lam_dev[3]);
int tg_adjoint_free(tg_plan *plan);

/* Seeded solenoidal initial field (this library's own generator; the
 * reference's Init_cond_energy_helicity_spectrum, init_cond_energy.cc:67-176,
 * is seeded by time(0)). kind 0: E(K) ~ K^4 exp(-2 (K/k0)^2) with Gaussian
 * amplitudes; kind 1: E(K) ~ K^(-5/3) on [kmin, kmax] with random phases. */
int tg_init_field(tg_plan *plan, int kind, double k0, double kmin, double kmax, double energy, uint64_t seed,
                  double *const V_dev[3]);

#ifdef __cplusplus
}
#endif
#endif
