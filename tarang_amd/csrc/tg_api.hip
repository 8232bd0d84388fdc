// tg_api.hip -- plan object and the extern "C" boundary (include/tarang_gpu.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/tarang_gpu.h"
#include "tg_adjoint.h"
#include "tg_chff.h"
#include "tg_diag.h"
#include "tg_kernels.h"
#include "tg_mhd.h"
#include "tg_pass.h"
#include "tg_sff.h"

using namespace tg;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(TG_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                         \
    } while (0)


}  // namespace

struct tg_plan {
    tg_config cfg{};
    Grid g{};
    hipStream_t stream = nullptr;
    int prec = TG_PREC_FP64;
    size_t elem = 16;     // bytes of one complex work element
    size_t nc = 0;        // complex modes per field
    size_t ncw = 0;       // complex entries of one padded work field
    int64_t dev_bytes = 0;
    // twiddle tables in the transform precision: x (Nx), y (Ny), m (Nz/2), n (Nz, k = 0..Nz/2)
    void *twx = nullptr, *twy = nullptr, *twm = nullptr, *twn = nullptr;
    void *U[6] = {};      // x/y-transformed velocity (0..2) and giver (3..5) lines
    void *P[9] = {};      // product spectra
    double *acc[3] = {};  // RK4 accumulator (tot_Vrhs, RK.cc:96); giver nonlinear term
    double *slots[2] = {};
    size_t nslots = 0;
    int slot_kb = -1;
    double *partial = nullptr;
    size_t partial_bytes = 0;
    double *small = nullptr;  // device scratch for reduction outputs / radii
    int red_blocks = 0;
    // tg_get_dt: one real field (Nx, Ny, Nz+2) and the three maxima, on first use
    double *rfield = nullptr, *umax_dev = nullptr;
    // random band forcing: per-mode supplies (e_k, h_k) of shells 0..n-1
    double *force_eh = nullptr;
    int force_eh_n = 0;
    double force_key[4] = {};  // eps, helicity, r_in, r_out of the table
    // chaining of RK4 steps: valid for the same state pointers and band
    bool unext_valid = false;
    double *unext_V[3] = {};
    Band unext_band{};
    // non-finite state detection: device flag set by the stage-4 kernels,
    // read back into pinned host memory at the end of every time-advance call
    int *nonfinite_dev = nullptr;
    int *nonfinite_host = nullptr;
    // host-side band config for the current step
    Band band{};
    int x_tile_columns = 0;  // fused x-pass tile width (0 = LDS-budget width)
    // SFF: quarter-wave twiddles exp(-i pi k / (2 Nx)), k < Nx
    void *twq = nullptr;
    // RBC buffers, allocated on first use: [0] theta accumulator (FP64 padded
    // work layout), [1..4] force inputs theta, u_x of two stage parities
    // (FP64 boundary layout)
    double *rbc[5] = {};
    // MHD buffers (FP64), allocated on first use: [0..2] accW (padded work
    // layout), [3..5] Z+, [6..8] Z- (boundary layout)
    double *mhd[9] = {};
    // adjoint: checkpoints (FP64 spectral, 3 components each), tape of the
    // real-space stage velocities (every x 4 stages x 3 components, packed
    // z lines in the transform precision), recompute state, cotangent and
    // accumulator (FP64 spectral)
    struct Adjoint {
        int nsteps = 0, every = 0, nsegs_used = 0, cached_seg = -1;
        int nseg = 0, tape_steps = 0;  // allocated capacity (checkpoints, tape steps)
        tg_fluid fluid{};
        double *ckpt = nullptr;
        void *tape = nullptr;
        double *vw = nullptr, *lg = nullptr, *acc = nullptr;
        bool recorded = false;
    } adj;
    // ChFF (TG_BASIS_CHFF) buffers, allocated on first use: the weighted copy
    // of an inverse transform's input, derivative/divergence and product
    // spectra, the nonlinear term (BDF1 right-hand sides), real velocity and
    // product arrays, Helmholtz scratch, the homogeneous solutions p+-,
    // u_x+- and the influence matrix (for the stored nu, dt)
    struct Chff {
        double *tmp = nullptr, *D = nullptr, *S = nullptr, *NL[3] = {}, *R[4] = {}, *X = nullptr, *Y = nullptr;
        double *pp = nullptr, *pm = nullptr, *vxp = nullptr, *vxm = nullptr, *M = nullptr;
        double infl_nu = -1.0, infl_dt = -1.0;
    } ch;
    // per-kernel HIP event timing (tg_profile_*)
    bool prof_on = false;
    struct ProfRec { int kid; hipEvent_t a, b; };
    std::vector<ProfRec> prof_recs;
    std::vector<hipEvent_t> prof_pool;
    double prof_ms[TG_KERNEL_COUNT] = {};
    int64_t prof_n[TG_KERNEL_COUNT] = {};
};

namespace {

hipEvent_t prof_event(tg_plan *p)
{
    if (!p->prof_pool.empty()) {
        hipEvent_t e = p->prof_pool.back();
        p->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one launch with events when profiling is on.
struct Prof {
    tg_plan *p;
    int kid;
    hipEvent_t a = nullptr;
    Prof(tg_plan *p_, int kid_) : p(p_), kid(kid_)
    {
        if (p->prof_on) {
            a = prof_event(p);
            if (a) (void)hipEventRecord(a, p->stream);
        }
    }
    ~Prof()
    {
        if (a) {
            hipEvent_t b = prof_event(p);
            if (b) {
                (void)hipEventRecord(b, p->stream);
                p->prof_recs.push_back({kid, a, b});
            } else {
                p->prof_pool.push_back(a);
            }
        }
    }
};

int prof_collect(tg_plan *p)
{
    if (p->prof_recs.empty()) return TG_OK;
    HIPCHK(hipStreamSynchronize(p->stream));
    for (auto &r : p->prof_recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            p->prof_ms[r.kid] += ms;
            p->prof_n[r.kid] += 1;
        }
        p->prof_pool.push_back(r.a);
        p->prof_pool.push_back(r.b);
    }
    p->prof_recs.clear();
    return TG_OK;
}

template <typename T>
int upload_twiddles(tg_plan *p, void **dst, int L, int count)
{
    std::vector<T> h(2 * (size_t)count);
    for (int k = 0; k < count; k++) {
        const double a = -2.0 * M_PI * (double)k / (double)L;
        h[2 * k] = (T)std::cos(a);
        h[2 * k + 1] = (T)std::sin(a);
    }
    HIPCHK(hipMalloc(dst, sizeof(T) * 2 * count));
    p->dev_bytes += sizeof(T) * 2 * count;
    HIPCHK(hipMemcpy(*dst, h.data(), sizeof(T) * 2 * count, hipMemcpyHostToDevice));
    return TG_OK;
}

int dmalloc(tg_plan *p, void **ptr, size_t bytes)
{
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(TG_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    p->dev_bytes += (int64_t)bytes;
    return TG_OK;
}

void free_plan(tg_plan *p)
{
    if (!p) return;
    void *ptrs[] = {p->twx, p->twy, p->twm, p->twn, p->twq, p->acc[0], p->acc[1], p->acc[2], p->slots[0],
                    p->slots[1], p->partial, p->small};
    for (void *q : p->rbc)
        if (q) (void)hipFree(q);
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (void *q : p->U)
        if (q) (void)hipFree(q);
    for (void *q : p->P)
        if (q) (void)hipFree(q);
    for (void *q : p->mhd)
        if (q) (void)hipFree(q);
    void *aptrs[] = {p->adj.ckpt, p->adj.tape, p->adj.vw, p->adj.lg, p->adj.acc, p->nonfinite_dev, p->rfield,
                     p->umax_dev, p->force_eh};
    for (void *q : aptrs)
        if (q) (void)hipFree(q);
    void *cptrs[] = {p->ch.tmp, p->ch.D, p->ch.S, p->ch.NL[0], p->ch.NL[1], p->ch.NL[2], p->ch.R[0], p->ch.R[1],
                     p->ch.R[2], p->ch.R[3], p->ch.X, p->ch.Y, p->ch.pp, p->ch.pm, p->ch.vxp, p->ch.vxm, p->ch.M};
    for (void *q : cptrs)
        if (q) (void)hipFree(q);
    if (p->nonfinite_host) (void)hipHostFree(p->nonfinite_host);
    for (auto &r : p->prof_recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : p->prof_pool) (void)hipEventDestroy(e);
    delete p;
}

int set_device(tg_plan *p) { HIPCHK(hipSetDevice(p->cfg.device)); return TG_OK; }

// Non-finite state detection around a time-advance call: the reference stops
// the run when the total energy is NaN (Ifluid_main.cc:127-128); here the
// stage-4 kernels flag any non-finite value they write into the state.
int nonfinite_begin(tg_plan *p)
{
    HIPCHK(launch_fill_int(p->nonfinite_dev, 1, 0, p->stream));
    return TG_OK;
}

int nonfinite_end(tg_plan *p, const char *what)
{
    HIPCHK(hipMemcpyAsync(p->nonfinite_host, p->nonfinite_dev, sizeof(int), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    if (*p->nonfinite_host)
        return fail(TG_ERR_NAN, "%s: the state holds non-finite values after the step (NaN or Inf)", what);
    return TG_OK;
}

int check_ptrs(const void *const *v, int n, const char *what)
{
    if (!v) return fail(TG_ERR_ARG, "%s: null pointer array", what);
    for (int i = 0; i < n; i++)
        if (!v[i]) return fail(TG_ERR_ARG, "%s: component %d is null", what, i);
    return TG_OK;
}

// Universal::Compute_modes_in_shell_array (universal_basic.cc:104-143): the
// modes of the full box, k_d in (-N_d/2, N_d/2], with ceil(K) = i, i <= imax
void modes_in_shell(const Grid &g, int imax, double *count)
{
    for (int i = 0; i <= imax; i++) count[i] = 0;
    const int n[3] = {g.nx, g.ny, g.nz};
    int lim[3];
    for (int d = 0; d < 3; d++) lim[d] = std::min((int)std::ceil(imax / g.kf[d]) + 1, n[d] / 2);
    for (int kx = -lim[0]; kx <= lim[0]; kx++)
        for (int ky = -lim[1]; ky <= lim[1]; ky++)
            for (int kz = -lim[2]; kz <= lim[2]; kz++) {
                if (kx <= -n[0] / 2 || ky <= -n[1] / 2 || kz <= -n[2] / 2) continue;
                const double a = kx * g.kf[0], b = ky * g.kf[1], c = kz * g.kf[2];
                const int idx = (int)std::ceil(std::sqrt(a * a + b * b + c * c));
                if (idx <= imax) count[idx] += 1;
            }
}

int ensure_slots(tg_plan *p, const tg_fluid *f)
{
    Band b{};
    if (f && f->force_kind == 1) {
        if (!(f->force_r_out > f->force_r_in) || f->force_r_in < 0)
            return fail(TG_ERR_ARG, "forcing band must satisfy 0 <= r_in < r_out");
        double kmin = std::fmin(p->g.kf[0], std::fmin(p->g.kf[1], p->g.kf[2]));
        int kb = (int)std::ceil(f->force_r_out / kmin);
        if (kb > p->g.nx / 2 - 1 || kb > p->g.ny / 2 - 1 || kb > p->g.nz / 2 - 1 || kb > 64)
            return fail(TG_ERR_ARG, "forcing band r_out=%g too large for this grid", f->force_r_out);
        b.kind = 1;
        b.kb = kb;
        b.r_in = f->force_r_in;
        b.r_out = f->force_r_out;
        b.eps = f->force_eps;
    } else if (f && f->force_kind == 2) {
        // random band forcing: per-mode supplies of the band's shells
        if (!(f->force_r_out > f->force_r_in) || f->force_r_in < 0)
            return fail(TG_ERR_ARG, "forcing band must satisfy 0 <= r_in < r_out");
        if (!(f->force_tau > 0) || !(f->force_eps >= 0))
            return fail(TG_ERR_ARG, "random forcing needs force_tau > 0 and force_eps >= 0");
        if (!(std::fabs(f->force_helicity) <= f->force_r_in * f->force_eps))
            return fail(TG_ERR_ARG, "random forcing needs |force_helicity| <= r_in force_eps");
        const int imax = (int)std::ceil(f->force_r_out), i0 = (int)std::floor(f->force_r_in) + 1;
        if (imax > 4096) return fail(TG_ERR_ARG, "forcing band r_out=%g too large", f->force_r_out);
        const double key[4] = {f->force_eps, f->force_helicity, f->force_r_in, f->force_r_out};
        if (!p->force_eh || std::memcmp(key, p->force_key, sizeof(key)) != 0) {
            // the per-mode supplies depend on the rates and the band only (not on the seed)
            const int nsh = imax - i0 + 1;
            std::vector<double> cnt(imax + 1, 0.0), eh(2 * (size_t)(imax + 1), 0.0);
            modes_in_shell(p->g, imax, cnt.data());
            for (int i = i0; i <= imax; i++)
                if (cnt[i] > 0) {
                    eh[2 * i] = f->force_eps / nsh / cnt[i];
                    eh[2 * i + 1] = f->force_helicity / nsh / cnt[i];
                }
            if (p->force_eh_n < imax + 1) {
                HIPCHK(hipStreamSynchronize(p->stream));
                if (p->force_eh) { (void)hipFree(p->force_eh); p->force_eh = nullptr; }
                int rc = dmalloc(p, (void **)&p->force_eh, eh.size() * sizeof(double));
                if (rc) return rc;
                p->force_eh_n = imax + 1;
            }
            HIPCHK(hipMemcpyAsync(p->force_eh, eh.data(), eh.size() * sizeof(double), hipMemcpyHostToDevice,
                                  p->stream));
            HIPCHK(hipStreamSynchronize(p->stream));
            std::memcpy(p->force_key, key, sizeof(key));
        }
        b.kind = 2;
        b.r_in = f->force_r_in;
        b.r_out = f->force_r_out;
        b.eps = f->force_eps;
        b.shell_eh = p->force_eh;
        b.tau = f->force_tau;
        b.seed = f->force_seed;
    } else if (f && f->force_kind != 0) {
        return fail(TG_ERR_ARG, "unknown force_kind %d", f->force_kind);
    }
    p->band = b;
    if (b.kind != 1) return TG_OK;
    if (p->slot_kb != b.kb) {
        // the stage-1 transform prepared for a continuation has no slots here
        p->unext_valid = false;
        for (int i = 0; i < 2; i++)
            if (p->slots[i]) { (void)hipFree(p->slots[i]); p->slots[i] = nullptr; }
        p->nslots = (size_t)band_nslots(b.kb);
        for (int i = 0; i < 2; i++) {
            int rc = dmalloc(p, (void **)&p->slots[i], p->nslots * 3 * 16);
            if (rc) return rc;
            HIPCHK(hipMemsetAsync(p->slots[i], 0, p->nslots * 3 * 16, p->stream));
        }
        p->slot_kb = b.kb;
    }
    return TG_OK;
}

// sincos: 0 Fourier (FFF), 1 sine, 2 cosine (SFF)
template <typename T>
int run_forward(tg_plan *p, const double *real_dev, double *spec_dev, int sincos)
{
    const double *R[1] = {real_dev};
    void *W[1] = {p->U[0]};
    { Prof pr(p, TG_KERNEL_Z_R2C); HIPCHK(launch_z_r2c<T>(p->g, R, 1, W, p->twm, p->twn, p->stream)); }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, -1, W, 1, p->twy, p->stream)); }
    double *O[1] = {spec_dev};
    if (p->g.basis == 1) {
        const double scale = 1.0 / (2.0 * p->g.nx * (double)p->g.ny * p->g.nz);
        Prof pr(p, TG_KERNEL_X_SC);
        HIPCHK(launch_xfwd_sc<T>(p->g, W, 1, sincos == TG_SC_SIN ? 1 : 0, O, scale, p->twx, p->twq, p->stream));
        // Chebyshev coefficients: a_0 = A_0, a_k = 2 A_k of the cosine transform
        if (sincos == TG_SC_CHEB) HIPCHK(launch_cheb_scale(p->g, spec_dev, spec_dev, 2.0, p->stream));
        return TG_OK;
    }
    const double scale = 1.0 / ((double)p->g.nx * p->g.ny * p->g.nz);
    { Prof pr(p, TG_KERNEL_X_FWD); HIPCHK(launch_xfwd_spec<T>(p->g, W, 1, O, scale, p->twx, p->stream)); }
    return TG_OK;
}

template <typename T>
int run_inverse(tg_plan *p, const double *spec_dev, double *real_dev, int sincos)
{
    double *V[1] = {const_cast<double *>(spec_dev)};
    void *W[1] = {p->U[0]};
    Band nb{};
    if (sincos == TG_SC_CHEB) {
        // the cosine coefficients A_0 = a_0, A_k = a_k / 2, into a copy (input preserved)
        HIPCHK(launch_cheb_scale(p->g, spec_dev, p->ch.tmp, 0.5, p->stream));
        V[0] = p->ch.tmp;
    }
    if (p->g.basis == 1) {
        Prof pr(p, TG_KERNEL_X_SC);
        HIPCHK(launch_xinv_sc<T>(p->g, V, 1, sincos == TG_SC_SIN ? 1 : 0, W, 0, p->twx, p->twq, p->stream));
    } else {
        Prof pr(p, TG_KERNEL_X_INV);
        HIPCHK(launch_xinv_state<T>(p->g, nb, V, 1, W, nullptr, 0, p->twx, p->stream));
    }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, +1, W, 1, p->twy, p->stream)); }
    double *R[1] = {real_dev};
    { Prof pr(p, TG_KERNEL_Z_C2R); HIPCHK(launch_z_c2r<T>(p->g, W, 1, R, p->twm, p->twn, p->stream)); }
    return TG_OK;
}

// Velocity x,y-inverse (U[0..2]) -> products -> y-forward (P[0..5])
template <typename T>
// trace_free: the nonlinear term is only used projected (an RK stage, or
// Compute_rhs), so the products are the five u_x^2 - u_z^2, u_y^2 - u_z^2,
// u_x u_y, u_y u_z, u_z u_x: the term less the gradient i K F[u_z^2], which
// the projection removes exactly, mode by mode.
int ns_middle(tg_plan *p, bool trace_free)
{
    const int np = trace_free ? 5 : 6;
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, +1, p->U, 3, p->twy, p->stream)); }
    {
        Prof pr(p, TG_KERNEL_Z_PRODUCTS);
        HIPCHK(launch_z_products<T>(p->g, p->U, p->P, trace_free ? 2 : 0, p->twm, p->twn, p->stream));
    }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, -1, p->P, np, p->twy, p->stream)); }
    return TG_OK;
}

// Hyperviscosity parameters: q = 0 (none) or 2..16 with nu_hyper >= 0;
// `allowed` false for the steps that take none (MHD, BDF1).
int check_hyper(const tg_fluid *f, bool allowed)
{
    if (f->hyper_exponent == 0) return TG_OK;
    if (!allowed) return fail(TG_ERR_ARG, "hyperviscosity (hyper_exponent %d) is taken by tg_rk4_step and the "
                                          "adjoint only", f->hyper_exponent);
    if (f->hyper_exponent < 2 || f->hyper_exponent > 16 || !(f->nu_hyper >= 0))
        return fail(TG_ERR_ARG, "need hyper_exponent 0 or 2..16 and nu_hyper >= 0");
    return TG_OK;
}

RkParams rk_params(const tg_plan *p, const tg_fluid *f, int stage)
{
    RkParams rk{};
    rk.stage = stage;
    rk.write_rhs = 0;
    const double nu = f ? f->nu : 0.0, dt = f ? f->dt : 0.0;
    rk.dt = dt;
    // Mult_field_exp_ksqr_dt factor = -dissipation_coefficient*a*dt (FluidVF.cc:342)
    rk.fh = -nu * 0.5 * dt;
    rk.f1 = -nu * 1.0 * dt;
    // and hyper_factor = -hyper_dissipation_coefficient*a*dt (FluidVF.cc:350-355)
    rk.hq = f ? f->hyper_exponent : 0;
    rk.fhh = f ? -f->nu_hyper * 0.5 * dt : 0.0;
    rk.f1h = f ? -f->nu_hyper * 1.0 * dt : 0.0;
    rk.scale = 1.0 / ((double)p->g.nx * p->g.ny * p->g.nz);
    return rk;
}

template <typename T>
int run_nlin(tg_plan *p, double *const V[3], double *const out[3], const tg_fluid *f, int write_rhs)
{
    if (write_rhs) {
        int rc = ensure_slots(p, f);
        if (rc) return rc;
    } else {
        p->band = Band{};
    }
    // Compute_nlin dealiases the state in place (compute_nlin.cc:61-62)
    if (p->g.dealias) HIPCHK(launch_dealias(p->g, V, 3, p->stream));
    HIPCHK(launch_xinv_state<T>(p->g, p->band, V, 3, p->U, p->slots[0], p->band.kind == 1 ? 2 : 0, p->twx, p->stream));
    int rc = ns_middle<T>(p, write_rhs != 0);
    if (rc) return rc;
    RkParams rk = rk_params(p, f, 0);
    rk.write_rhs = write_rhs;
    rk.trace_free = write_rhs != 0;
    HIPCHK(launch_x_nlin<T>(p->g, p->band, rk, p->P, nullptr, nullptr, nullptr, p->slots[0], nullptr, out, p->twx,
                            p->stream));
    p->unext_valid = false;
    return TG_OK;
}

// N_i = FT[d_j(x_j w_i)] with x the advecting field X and w = Y restricted
// by the x-inverse filter flags (0 none; 8 r_lo < K <= r_hi, Fill_array_shell,
// universal_ET.cc:880-892; 16 ring and 32 cylindrical ring about the axis
// dirn, Fill_array_ring / Fill_array_cylindrical_ring, :897-933).
template <typename T>
// receiver_ready: U[0..2] still hold the x- and y-inverse of X from the
// previous call (the z pass only reads them), so only the giver is
// transformed (consecutive evaluations with one receiver: flux spheres,
// shell-to-shell givers).
int run_giver_nlin(tg_plan *p, double *const X[3], double *const Y[3], int filter, double r_lo, double r_hi,
                   double *const out[3], bool receiver_ready, double a_lo, double a_hi, int dirn)
{
    Band nb{};
    {
        Prof pr(p, TG_KERNEL_X_INV);
        if (!receiver_ready) HIPCHK(launch_xinv_state<T>(p->g, nb, X, 3, p->U, nullptr, 0, p->twx, p->stream));
        HIPCHK(launch_xinv_state<T>(p->g, nb, Y, 3, p->U + 3, nullptr, filter, p->twx, p->stream, r_lo, r_hi,
                                    a_lo, a_hi, dirn));
    }
    {
        Prof pr(p, TG_KERNEL_Y_FFT);
        if (receiver_ready) HIPCHK(launch_colfft<T>(p->g, 1, +1, p->U + 3, 3, p->twy, p->stream));
        else HIPCHK(launch_colfft<T>(p->g, 1, +1, p->U, 6, p->twy, p->stream));
    }
    { Prof pr(p, TG_KERNEL_Z_PRODUCTS); HIPCHK(launch_z_products<T>(p->g, p->U, p->P, 1, p->twm, p->twn, p->stream)); }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, -1, p->P, 9, p->twy, p->stream)); }
    RkParams rk = rk_params(p, nullptr, -1);
    Prof pr(p, TG_KERNEL_X_NLIN_GIVER);
    HIPCHK(launch_x_nlin<T>(p->g, nb, rk, p->P, nullptr, nullptr, nullptr, nullptr, nullptr, out, p->twx,
                            p->stream));
    p->unext_valid = false;
    return TG_OK;
}

template <typename T>
int run_rk4(tg_plan *p, double *const V[3], const tg_fluid *f, int nsteps, int flags)
{
    int rc = ensure_slots(p, f);
    if (rc) return rc;
    const Band &ob = p->unext_band, &nb = p->band;
    const bool same_band = ob.kind == nb.kind && ob.kb == nb.kb && ob.r_in == nb.r_in && ob.r_out == nb.r_out &&
                           ob.eps == nb.eps;
    const bool cont = (flags & TG_STEP_CONTINUE) && p->unext_valid && same_band && p->unext_V[0] == V[0] &&
                      p->unext_V[1] == V[1] && p->unext_V[2] == V[2];
    rc = nonfinite_begin(p);
    if (rc) return rc;
    if (!cont) {
        Prof pr(p, TG_KERNEL_X_INV);
        HIPCHK(launch_xinv_state<T>(p->g, p->band, V, 3, p->U, p->slots[0], 1 | (p->band.kind == 1 ? 2 : 0), p->twx,
                                    p->stream));
    }
    for (int step = 0; step < nsteps; step++) {
        for (int s = 1; s <= 4; s++) {
            rc = ns_middle<T>(p, true);
            if (rc) return rc;
            RkParams rk = rk_params(p, f, s);
            rk.trace_free = 1;
            if (s == 4) rk.nonfinite = p->nonfinite_dev;
            Prof pr(p, TG_KERNEL_X_NLIN_RK);
            HIPCHK(launch_x_nlin<T>(p->g, p->band, rk, p->P, V, p->acc, p->U, p->slots[(s - 1) & 1],
                                    p->slots[s & 1], nullptr, p->twx, p->stream, p->x_tile_columns));
        }
        if (p->cfg.strong_reality) {
            Prof pr(p, TG_KERNEL_REALITY);
            HIPCHK(launch_reality(p->g, V, 3, p->stream));
            HIPCHK(launch_xinv_state<T>(p->g, p->band, V, 3, p->U, p->slots[0], 1 | 4 | (p->band.kind == 1 ? 2 : 0),
                                        p->twx, p->stream));
        }
    }
    p->unext_valid = true;
    p->unext_band = p->band;
    for (int c = 0; c < 3; c++) p->unext_V[c] = V[c];
    return nonfinite_end(p, "tg_rk4_step");
}

int ensure_mhd(tg_plan *p)
{
    for (int i = 0; i < 9; i++)
        if (!p->mhd[i]) {
            int rc = dmalloc(p, (void **)&p->mhd[i], (i < 3 ? p->ncw : p->nc) * 16);
            if (rc) return rc;
        }
    return TG_OK;
}

// N+ (giver form) and N- (transposed) of the Elsasser fields whose x-inverse
// sits in U[0..2] (Z-) and U[3..5] (Z+).
template <typename T>
int mhd_nonlinear(tg_plan *p, double *const *Np, double *const *Nm)
{
    Band nb{};
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, +1, p->U, 6, p->twy, p->stream)); }
    { Prof pr(p, TG_KERNEL_Z_PRODUCTS); HIPCHK(launch_z_products<T>(p->g, p->U, p->P, 1, p->twm, p->twn, p->stream)); }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, -1, p->P, 9, p->twy, p->stream)); }
    Prof pr(p, TG_KERNEL_X_NLIN_GIVER);
    RkParams rk = rk_params(p, nullptr, -1);
    HIPCHK(launch_x_nlin<T>(p->g, nb, rk, p->P, nullptr, nullptr, nullptr, nullptr, nullptr, Np, p->twx, p->stream));
    rk.stage = -2;
    HIPCHK(launch_x_nlin<T>(p->g, nb, rk, p->P, nullptr, nullptr, nullptr, nullptr, nullptr, Nm, p->twx, p->stream));
    return TG_OK;
}

template <typename T>
int mhd_xinv(tg_plan *p, double *const *Zp, double *const *Zm)
{
    Band nb{};
    Prof pr(p, TG_KERNEL_X_INV);
    HIPCHK(launch_xinv_state<T>(p->g, nb, Zm, 3, p->U, nullptr, 0, p->twx, p->stream));
    HIPCHK(launch_xinv_state<T>(p->g, nb, Zp, 3, p->U + 3, nullptr, 0, p->twx, p->stream));
    return TG_OK;
}

template <typename T>
int run_rk4_mhd(tg_plan *p, double *const V[3], double *const W[3], const tg_fluid *f, double eta, int nsteps)
{
    int rc = ensure_mhd(p);
    if (rc) return rc;
    double *accW[3] = {p->mhd[0], p->mhd[1], p->mhd[2]};
    double *Zp[3] = {p->mhd[3], p->mhd[4], p->mhd[5]};
    double *Zm[3] = {p->mhd[6], p->mhd[7], p->mhd[8]};
    p->unext_valid = false;
    rc = nonfinite_begin(p);
    if (rc) return rc;
    // stage-1 input of the first step; the stage-4 x pass of each step leaves
    // that of the next
    HIPCHK(launch_mhd_elsasser(p->g, V, W, Zp, Zm, p->stream));
    rc = mhd_xinv<T>(p, Zp, Zm);
    if (rc) return rc;
    for (int step = 0; step < nsteps; step++) {
        for (int s = 1; s <= 4; s++) {
            { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, +1, p->U, 6, p->twy, p->stream)); }
            {
                // the eight trace-shifted products (P_zz enters N_u only as a gradient)
                Prof pr(p, TG_KERNEL_Z_PRODUCTS);
                HIPCHK(launch_z_products<T>(p->g, p->U, p->P, 3, p->twm, p->twn, p->stream));
            }
            { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, -1, p->P, 8, p->twy, p->stream)); }
            // one x pass: both nonlinear terms, the stage for u and b, the next Z+- x-inverse
            XmhdArgs xa{};
            for (int i = 0; i < 9; i++) xa.P[i] = p->P[i];
            for (int c = 0; c < 3; c++) {
                xa.V[c] = V[c];
                xa.W[c] = W[c];
                xa.accV[c] = p->acc[c];
                xa.accW[c] = accW[c];
            }
            for (int i = 0; i < 6; i++) xa.U[i] = p->U[i];
            xa.tw = p->twx;
            xa.g = p->g;
            xa.stage = s;
            xa.dt = f->dt;
            xa.fh_u = -f->nu * 0.5 * f->dt;
            xa.f1_u = -f->nu * 1.0 * f->dt;
            xa.fh_b = -eta * 0.5 * f->dt;
            xa.f1_b = -eta * 1.0 * f->dt;
            xa.scale = 1.0 / ((double)p->g.nx * p->g.ny * p->g.nz);
            xa.nonfinite = s == 4 ? p->nonfinite_dev : nullptr;
            Prof pr(p, TG_KERNEL_X_NLIN_RK);
            HIPCHK(launch_x_mhd<T>(xa, p->stream));
        }
        if (p->cfg.strong_reality) {
            Prof pr(p, TG_KERNEL_REALITY);
            HIPCHK(launch_reality(p->g, V, 3, p->stream));
            HIPCHK(launch_reality(p->g, W, 3, p->stream));
            if (step + 1 < nsteps) {
                // the next stage-1 input on the two planes the condition rewrote
                Band nb{};
                HIPCHK(launch_mhd_elsasser_planes(p->g, V, W, Zp, Zm, p->stream));
                HIPCHK(launch_xinv_state<T>(p->g, nb, Zm, 3, p->U, nullptr, 4, p->twx, p->stream));
                HIPCHK(launch_xinv_state<T>(p->g, nb, Zp, 3, p->U + 3, nullptr, 4, p->twx, p->stream));
            }
        }
    }
    return nonfinite_end(p, "tg_rk4_step_mhd");
}

template <typename T>
int run_nlin_mhd(tg_plan *p, double *const V[3], double *const W[3], double *const NU[3], double *const NW[3])
{
    int rc = ensure_mhd(p);
    if (rc) return rc;
    double *Zp[3] = {p->mhd[3], p->mhd[4], p->mhd[5]};
    double *Zm[3] = {p->mhd[6], p->mhd[7], p->mhd[8]};
    p->unext_valid = false;
    if (p->g.dealias) {
        HIPCHK(launch_dealias(p->g, V, 3, p->stream));
        HIPCHK(launch_dealias(p->g, W, 3, p->stream));
    }
    HIPCHK(launch_mhd_elsasser(p->g, V, W, Zp, Zm, p->stream));
    rc = mhd_xinv<T>(p, Zp, Zm);
    if (!rc) rc = mhd_nonlinear<T>(p, NU, NW);
    if (rc) return rc;
    HIPCHK(launch_mhd_combine(p->g, NU, NW, p->stream));
    return TG_OK;
}

// ---------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------
size_t tape_field_elems(const tg_plan *p) { return (size_t)p->g.nx * p->g.ny * (p->g.nz / 2); }

void *tape_ptr(tg_plan *p, int step, int stage, int comp)
{
    const size_t f = tape_field_elems(p) * p->elem;
    return (char *)p->adj.tape + ((size_t)(step * 4 + stage) * 3 + comp) * f;
}

void adj_free(tg_plan *p)
{
    auto &a = p->adj;
    const size_t nc = p->nc;
    void *ptrs[] = {a.ckpt, a.tape, a.vw, a.lg, a.acc};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p->dev_bytes -= (int64_t)((size_t)a.nseg * 3 * nc * 16 + (size_t)a.tape_steps * 12 * tape_field_elems(p) * p->elem +
                              (a.vw ? 9 * nc * 16 : 0));
    a = tg_plan::Adjoint{};
}

// Adjoint buffers for at least nseg checkpoints and tape_steps steps of tape
// (kept when the current ones are large enough).
int adj_alloc(tg_plan *p, int nseg, int tape_steps)
{
    auto &a = p->adj;
    if (a.vw && a.nseg >= nseg && a.tape_steps >= tape_steps) return TG_OK;
    if (a.vw) {
        nseg = std::max(nseg, a.nseg);
        tape_steps = std::max(tape_steps, a.tape_steps);
    }
    (void)hipStreamSynchronize(p->stream);
    adj_free(p);
    const size_t nc = p->nc;
    int rc = dmalloc(p, (void **)&a.ckpt, (size_t)nseg * 3 * nc * 16);
    if (!rc) rc = dmalloc(p, &a.tape, (size_t)tape_steps * 12 * tape_field_elems(p) * p->elem);
    if (!rc) rc = dmalloc(p, (void **)&a.vw, 3 * nc * 16);
    if (!rc) rc = dmalloc(p, (void **)&a.lg, 3 * nc * 16);
    if (!rc) rc = dmalloc(p, (void **)&a.acc, 3 * nc * 16);
    a.nseg = nseg;
    a.tape_steps = tape_steps;
    if (rc) {
        adj_free(p);
        return rc;
    }
    return TG_OK;
}

// nsteps unforced RK4 steps of V (no reality condition), storing the
// real-space velocity of every stage input in tape step slots 0..nsteps-1.
template <typename T>
int record_steps(tg_plan *p, double *const V[3], const tg_fluid *f, int nsteps)
{
    Band nb{};
    p->band = nb;
    {
        Prof pr(p, TG_KERNEL_X_INV);
        HIPCHK(launch_xinv_state<T>(p->g, nb, V, 3, p->U, nullptr, 1, p->twx, p->stream));
    }
    for (int step = 0; step < nsteps; step++)
        for (int s = 1; s <= 4; s++) {
            void *tp[3] = {tape_ptr(p, step, s - 1, 0), tape_ptr(p, step, s - 1, 1), tape_ptr(p, step, s - 1, 2)};
            { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, +1, p->U, 3, p->twy, p->stream)); }
            {
                Prof pr(p, TG_KERNEL_Z_PRODUCTS);
                HIPCHK(launch_z_products<T>(p->g, p->U, p->P, 2, p->twm, p->twn, p->stream, tp));
            }
            { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, -1, p->P, 5, p->twy, p->stream)); }
            RkParams rk = rk_params(p, f, s);
            rk.trace_free = 1;
            Prof pr(p, TG_KERNEL_X_NLIN_RK);
            HIPCHK(launch_x_nlin<T>(p->g, nb, rk, p->P, V, p->acc, p->U, nullptr, nullptr, nullptr, p->twx,
                                    p->stream, p->x_tile_columns));
        }
    p->unext_valid = false;
    return TG_OK;
}

// One transposed right-hand side application with the stage epilogue:
// cotangent G, taped velocity tp, outputs per AdjParams.
template <typename T>
// One backward stage.  sym_done: the symmetrised fields of this stage are
// already in S (made by the previous stage's fused x pass); fuse_next: this
// stage's x pass also makes the next stage's symmetrised fields (the next
// stage then skips launch_xinv_sym and its cotangent never goes to HBM).
int adj_stage(tg_plan *p, double *const G[3], void *const tp[3], const AdjParams &ap, double *const lam[3],
              double *const acc[3], double *const lg[3], bool sym_done = false, bool fuse_next = false)
{
    void *S[5] = {p->P[0], p->P[1], p->P[2], p->P[3], p->P[4]};
    void *Q[3] = {p->P[6], p->P[7], p->P[8]};
    if (!sym_done) { Prof pr(p, TG_KERNEL_X_INV_SYM); HIPCHK(launch_xinv_sym<T>(p->g, G, S, p->twx, p->stream)); }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, +1, S, 5, p->twy, p->stream)); }
    { Prof pr(p, TG_KERNEL_Z_ADJ); HIPCHK(launch_z_adj<T>(p->g, S, Q, tp, p->twm, p->twn, p->stream)); }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, -1, Q, 3, p->twy, p->stream)); }
    Prof pr(p, TG_KERNEL_X_FWD_ADJ);
    HIPCHK(launch_xfwd_adj<T>(p->g, ap, Q, lam, acc, lg, p->twx, p->stream, fuse_next ? S : nullptr));
    return TG_OK;
}

AdjParams adj_params(const tg_plan *p, const tg_fluid *f, int mode)
{
    AdjParams ap{};
    ap.mode = mode;
    ap.dt = f ? f->dt : 0.0;
    ap.fh = f ? -f->nu * 0.5 * f->dt : 0.0;
    ap.f1 = f ? -f->nu * 1.0 * f->dt : 0.0;
    ap.hq = f ? f->hyper_exponent : 0;
    ap.fhh = f ? -f->nu_hyper * 0.5 * f->dt : 0.0;
    ap.f1h = f ? -f->nu_hyper * 1.0 * f->dt : 0.0;
    ap.scale = 1.0 / ((double)p->g.nx * p->g.ny * p->g.nz);
    return ap;
}

// lam <- (dV_{n+1}/dV_n)^T lam for the taped step slot `step`.
template <typename T>
// first: this step's first stage makes its own symmetrised fields; last: no
// later stage follows within the segment (the recomputation of the previous
// segment reuses the S buffers), so the last stage does not fuse.
int adj_step(tg_plan *p, int step, double *const lam[3], bool first, bool last)
{
    const size_t nc = p->nc;
    double *lg[3] = {p->adj.lg, p->adj.lg + 2 * nc, p->adj.lg + 4 * nc};
    double *acc[3] = {p->adj.acc, p->adj.acc + 2 * nc, p->adj.acc + 4 * nc};
    const tg_fluid *f = &p->adj.fluid;
    for (int s = 3; s >= 0; s--) {
        void *tp[3] = {tape_ptr(p, step, s, 0), tape_ptr(p, step, s, 1), tape_ptr(p, step, s, 2)};
        const bool sym_done = !(s == 3 && first);
        const bool fuse_next = !(s == 0 && last);
        int rc = adj_stage<T>(p, s == 3 ? lam : lg, tp, adj_params(p, f, s), lam, acc, lg, sym_done, fuse_next);
        if (rc) return rc;
    }
    return TG_OK;
}

int seg_steps(const tg_plan *p, int seg)
{
    const int left = p->adj.nsteps - seg * p->adj.every;
    return left < p->adj.every ? left : p->adj.every;
}

template <typename T>
int run_adjoint_forward(tg_plan *p, double *const V[3], const tg_fluid *f, int nsteps, int every)
{
    if (every > nsteps) every = nsteps;
    const int nseg = (nsteps + every - 1) / every;
    int rc = adj_alloc(p, nseg, every);
    if (rc) return rc;
    auto &a = p->adj;
    a.nsteps = nsteps;
    a.every = every;
    a.nsegs_used = nseg;
    a.fluid = *f;
    a.recorded = false;
    const size_t nc = p->nc;
    for (int seg = 0; seg < nseg; seg++) {
        for (int c = 0; c < 3; c++)
            HIPCHK(hipMemcpyAsync(a.ckpt + ((size_t)seg * 3 + c) * 2 * nc, V[c], nc * 16, hipMemcpyDeviceToDevice,
                                  p->stream));
        const int n = seg_steps(p, seg);
        if (seg == nseg - 1) rc = record_steps<T>(p, V, f, n);
        else rc = run_rk4<T>(p, V, f, n, 0);
        if (rc) return rc;
    }
    a.cached_seg = nseg - 1;
    a.recorded = true;
    p->unext_valid = false;
    return TG_OK;
}

template <typename T>
int run_adjoint_backward(tg_plan *p, double *const lam[3])
{
    auto &a = p->adj;
    const size_t nc = p->nc;
    for (int seg = a.nsegs_used - 1; seg >= 0; seg--) {
        const int n = seg_steps(p, seg);
        if (a.cached_seg != seg) {
            double *vw[3] = {a.vw, a.vw + 2 * nc, a.vw + 4 * nc};
            for (int c = 0; c < 3; c++)
                HIPCHK(hipMemcpyAsync(vw[c], a.ckpt + ((size_t)seg * 3 + c) * 2 * nc, nc * 16,
                                      hipMemcpyDeviceToDevice, p->stream));
            int rc = record_steps<T>(p, vw, &a.fluid, n);
            if (rc) return rc;
            a.cached_seg = seg;
        }
        for (int st = n - 1; st >= 0; st--) {
            int rc = adj_step<T>(p, st, lam, st == n - 1, st == 0);
            if (rc) return rc;
        }
    }
    return TG_OK;
}

template <typename T>
int run_rhs_vjp(tg_plan *p, double *const V[3], double *const lam[3], double *const out[3])
{
    // one tape step slot for u = I(D V)
    if (!p->adj.tape) {
        int rc = adj_alloc(p, 1, 1);
        if (rc) return rc;
        p->adj.recorded = false;
    }
    Band nb{};
    void *tp[3] = {tape_ptr(p, 0, 0, 0), tape_ptr(p, 0, 0, 1), tape_ptr(p, 0, 0, 2)};
    { Prof pr(p, TG_KERNEL_X_INV); HIPCHK(launch_xinv_state<T>(p->g, nb, V, 3, p->U, nullptr, 1, p->twx, p->stream)); }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, +1, p->U, 3, p->twy, p->stream)); }
    {
        Prof pr(p, TG_KERNEL_Z_PRODUCTS);
        HIPCHK(launch_z_products<T>(p->g, p->U, p->P, 0, p->twm, p->twn, p->stream, tp));
    }
    p->unext_valid = false;
    // the tape slot now differs from any recorded segment
    if (p->adj.cached_seg >= 0) p->adj.cached_seg = -1;
    return adj_stage<T>(p, lam, tp, adj_params(p, nullptr, ADJ_RHS), nullptr, nullptr, out);
}

// ---------------------------------------------------------------------------
// Rayleigh-Benard convection on the SFF basis
// ---------------------------------------------------------------------------
int ensure_rbc(tg_plan *p)
{
    if (!p->rbc[0]) {
        int rc = dmalloc(p, (void **)&p->rbc[0], p->ncw * 16);
        if (rc) return rc;
    }
    for (int i = 1; i < 5; i++)
        if (!p->rbc[i]) {
            int rc = dmalloc(p, (void **)&p->rbc[i], p->nc * 16);
            if (rc) return rc;
        }
    return TG_OK;
}

RbcParams rbc_params(const tg_plan *p, const tg_rbc *r, int stage)
{
    RbcParams rp{};
    rp.stage = stage;
    const double nu = r ? std::sqrt(r->Pr / r->Ra) : 0.0, kappa = r ? 1.0 / std::sqrt(r->Pr * r->Ra) : 0.0;
    const double dt = r ? r->dt : 0.0;
    rp.dt = dt;
    rp.fh_u = -nu * 0.5 * dt;
    rp.f1_u = -nu * 1.0 * dt;
    rp.fh_t = -kappa * 0.5 * dt;
    rp.f1_t = -kappa * 1.0 * dt;
    rp.scale = 1.0 / (2.0 * p->g.nx * (double)p->g.ny * p->g.nz);
    return rp;
}

// y-inverse of U[0..3], z products, y-forward of P[0..8]
template <typename T>
// shift: an RK stage, whose velocity products are trace-shifted (xx - zz,
// yy - zz; P[2], zz, neither formed nor transformed)
int rbc_middle(tg_plan *p, bool shift)
{
    void *F[9];
    int nf = 0;
    for (int i = 0; i < 9; i++)
        if (!(shift && i == 2)) F[nf++] = p->P[i];
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, +1, p->U, 4, p->twy, p->stream)); }
    { Prof pr(p, TG_KERNEL_Z_RBC); HIPCHK(launch_z_rbc<T>(p->g, p->U, p->P, p->twm, p->twn, p->stream, shift)); }
    { Prof pr(p, TG_KERNEL_Y_FFT); HIPCHK(launch_colfft<T>(p->g, 1, -1, F, nf, p->twy, p->stream)); }
    return TG_OK;
}

template <typename T>
int run_nlin_rbc(tg_plan *p, double *const V[3], double *T_, double *const out[3], double *outT, const tg_rbc *r)
{
    int rc = ensure_rbc(p);
    if (rc) return rc;
    double *S[4] = {V[0], V[1], V[2], T_};
    const double *f_in[2] = {p->rbc[1], p->rbc[2]};
    if (r) {
        // the force sees the fields before Compute_nlin dealiases them
        HIPCHK(hipMemcpyAsync(p->rbc[1], T_, p->nc * 16, hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(p->rbc[2], V[0], p->nc * 16, hipMemcpyDeviceToDevice, p->stream));
    }
    if (p->g.dealias) {
        HIPCHK(launch_dealias(p->g, S, 3, p->stream));
        HIPCHK(launch_dealias(p->g, S + 3, 1, p->stream));
    }
    {
        Prof pr(p, TG_KERNEL_X_SC);
        HIPCHK(launch_xinv_sc<T>(p->g, S, 4, kRbcStateParity, p->U, 0, p->twx, p->twq, p->stream));
    }
    rc = rbc_middle<T>(p, false);
    if (rc) return rc;
    RbcParams rp = rbc_params(p, r, 0);
    rp.write_rhs = r ? 1 : 0;
    double *O[4] = {out[0], out[1], out[2], outT};
    Prof pr(p, TG_KERNEL_X_RBC);
    HIPCHK(launch_x_rbc<T>(p->g, rp, p->P, nullptr, nullptr, nullptr, f_in, nullptr, O, p->twx, p->twq, p->stream));
    p->unext_valid = false;
    return TG_OK;
}

template <typename T>
int run_rk4_rbc(tg_plan *p, double *const V[3], double *T_, const tg_rbc *r, int nsteps)
{
    int rc = ensure_rbc(p);
    if (rc) return rc;
    double *S[4] = {V[0], V[1], V[2], T_};
    double *acc[4] = {p->acc[0], p->acc[1], p->acc[2], p->rbc[0]};
    const int reality = p->cfg.strong_reality;
    p->unext_valid = false;
    rc = nonfinite_begin(p);
    if (rc) return rc;
    {
        Prof pr(p, TG_KERNEL_X_SC);
        HIPCHK(launch_xinv_sc<T>(p->g, S, 4, kRbcStateParity, p->U, 1, p->twx, p->twq, p->stream));
    }
    for (int step = 0; step < nsteps; step++) {
        for (int s = 1; s <= 4; s++) {
            rc = rbc_middle<T>(p, true);
            if (rc) return rc;
            // force of stage s: the state (stage 1) or the undealiased stage
            // input written by stage s-1, ping-pong by stage parity
            const int rd = (s - 1) & 1, wr = s & 1;
            const double *f_in[2] = {s == 1 ? T_ : p->rbc[1 + 2 * rd], s == 1 ? V[0] : p->rbc[2 + 2 * rd]};
            double *f_out[2] = {p->rbc[1 + 2 * wr], p->rbc[2 + 2 * wr]};
            RbcParams rp = rbc_params(p, r, s);
            rp.trace_free = 1;
            if (s == 4) rp.nonfinite = p->nonfinite_dev;
            Prof pr(p, TG_KERNEL_X_RBC);
            HIPCHK(launch_x_rbc<T>(p->g, rp, p->P, S, acc, p->U, f_in, f_out, nullptr, p->twx, p->twq, p->stream));
        }
        if (reality == 1 || reality == 2) {
            Prof pr(p, TG_KERNEL_REALITY);
            HIPCHK(launch_sff_reality(p->g, S, 4, reality, p->stream));
            // the next stage-1 input: redo the x-inverse of the kz = 0 and Nz/2 columns
            HIPCHK(launch_xinv_sc<T>(p->g, S, 4, kRbcStateParity, p->U, 1 | 4, p->twx, p->twq, p->stream));
        }
    }
    return nonfinite_end(p, "tg_rk4_step_rbc");
}

int reduce_to_host(tg_plan *p, int mode, const double *const *A, const double *const *B, int ncomp, int n,
                   const double *radii_host, int nbins, int nq, int kcut, double r_lo, double r_hi, double *out)
{
    const int width = nq * nbins;
    double *radii_dev = nullptr;
    size_t need_small = sizeof(double) * (size_t)(width + nbins + 8);
    if (need_small > 65536) return fail(TG_ERR_ARG, "too many bins (%d)", nbins);
    double *outd = p->small;
    if (radii_host) {
        radii_dev = p->small + width + 4;
        HIPCHK(hipMemcpyAsync(radii_dev, radii_host, sizeof(double) * nbins, hipMemcpyHostToDevice, p->stream));
    }
    const size_t need = (size_t)p->red_blocks * width * sizeof(double);
    if (need > p->partial_bytes) return fail(TG_ERR_ARG, "reduction needs %zu bytes of partials", need);
    HIPCHK(launch_reduce(p->g, mode, A, B, ncomp, n, radii_dev, nbins, nq, kcut, r_lo, r_hi, p->partial,
                         p->red_blocks, outd, p->stream));
    HIPCHK(hipMemcpyAsync(out, outd, sizeof(double) * width, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return TG_OK;
}

int max_radius_inside(const Grid &g)
{
    double k = std::fmin((g.nx / 2) * g.kf[0], (g.ny / 2) * g.kf[1]);
    k = std::fmin(k, (g.nz / 2) * g.kf[2]);
    return (int)k;
}

int min_radius_outside(const Grid &g)
{
    double a = (g.nx / 2) * g.kf[0], b = (g.ny / 2) * g.kf[1], c = (g.nz / 2) * g.kf[2];
    return (int)std::ceil(std::sqrt(a * a + b * b + c * c));
}

// Anis_max_Krho_radius_inside (FFF_pencil_inline.h:784-798)
int aniso_max_krho_inside(const Grid &g, int dirn)
{
    const double a = (g.nx / 2) * g.kf[0], b = (g.ny / 2) * g.kf[1], c = (g.nz / 2) * g.kf[2];
    const double k = dirn == 1 ? std::fmin(b, c) : (dirn == 2 ? std::fmin(a, c) : std::fmin(a, b));
    return (int)k;
}

// Anisotropic binned reduction into out[q * nbins_total + bin], in windows of
// bins that fit the per-wavefront LDS bins and the partial buffer.  radii
// (nshell + 1 entries, may be null) and aux (nsub + 1 entries) are host
// arrays, uploaded behind the output window in the plan's small buffer.
int reduce_aniso_to_host(tg_plan *p, int mode, const double *const *A, const double *const *B, int ncomp, int n,
                         int nq, const double *radii_host, int nshell, const double *aux_host, int nsub, int dirn,
                         int nbins_total, int kcut, double *out)
{
    const int win = std::max(1, std::min(nbins_total, std::min(1100, 2200 / nq)));
    const size_t off_aux = (size_t)nq * win + 8, off_radii = off_aux + (size_t)nsub + 8;
    const size_t need = sizeof(double) * (off_radii + (size_t)nshell + 8);
    if (need > 65536) return fail(TG_ERR_ARG, "too many shells or sub-bins (%d, %d)", nshell, nsub);
    double *outd = p->small, *auxd = p->small + off_aux, *radd = radii_host ? p->small + off_radii : nullptr;
    HIPCHK(hipMemcpyAsync(auxd, aux_host, sizeof(double) * (nsub + 1), hipMemcpyHostToDevice, p->stream));
    if (radii_host)
        HIPCHK(hipMemcpyAsync(radd, radii_host, sizeof(double) * (nshell + 1), hipMemcpyHostToDevice, p->stream));
    std::vector<double> tmp((size_t)nq * win);
    for (int lo = 0; lo < nbins_total; lo += win) {
        const int nb = std::min(win, nbins_total - lo);
        RedAniso an{auxd, nsub, nshell, dirn, lo};
        HIPCHK(launch_reduce(p->g, mode, A, B, ncomp, n, radd, nb, nq, kcut, 0.0, 0.0, p->partial, p->red_blocks,
                             outd, p->stream, &an));
        HIPCHK(hipMemcpyAsync(tmp.data(), outd, sizeof(double) * nq * nb, hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
        for (int q = 0; q < nq; q++)
            for (int b = 0; b < nb; b++) out[(size_t)q * nbins_total + lo + b] = tmp[(size_t)q * nb + b];
    }
    return TG_OK;
}

int check_edges(const double *e, int n, const char *what)
{
    if (!e || n < 1) return fail(TG_ERR_ARG, "%s: need at least one interval", what);
    for (int i = 1; i <= n; i++)
        if (!(e[i] >= e[i - 1])) return fail(TG_ERR_ARG, "%s must not decrease", what);
    return TG_OK;
}

}  // namespace

#define DISPATCH_PREC(p, CALL) ((p)->prec == TG_PREC_FP32 ? CALL<float> : CALL<double>)

extern "C" {

const char *tg_last_error_string(void) { return g_last_error.c_str(); }

const char *tg_version(void) { return "tarang_amd 0.1 (gfx950)"; }

int tg_plan_create(const tg_config *cfg, tg_plan **plan)
{
    if (!cfg || !plan) return fail(TG_ERR_ARG, "null argument");
    *plan = nullptr;
    if (cfg->basis != TG_BASIS_FFF && cfg->basis != TG_BASIS_SFF && cfg->basis != TG_BASIS_CHFF)
        return fail(TG_ERR_ARG, "unsupported basis %d", cfg->basis);
    if (cfg->basis == TG_BASIS_CHFF && (cfg->precision != TG_PREC_FP64 || cfg->strong_reality != 0))
        return fail(TG_ERR_ARG, "ChFF plans are FP64 with strong_reality 0");
    if (cfg->strong_reality < 0 || cfg->strong_reality > (cfg->basis == TG_BASIS_SFF ? 2 : 1))
        return fail(TG_ERR_ARG, "strong_reality %d out of range for this basis", cfg->strong_reality);
    if (cfg->precision != TG_PREC_FP64 && cfg->precision != TG_PREC_FP32)
        return fail(TG_ERR_ARG, "unknown precision %d", cfg->precision);
    {
        // transform lengths with compiled kernels: Nx, Ny in 8..1024 and
        // Nz/2 in 8..512 of the form 2^a, and in FP64 also 3, 5 or 15 times
        // a power of two (multiples of 8: 24, 40, 48, ..., 960)
        const bool fp32 = cfg->precision == TG_PREC_FP32;
        if (!l_compiled(cfg->nx, fp32) || !l_compiled(cfg->ny, fp32) || cfg->nz % 2 != 0 ||
            !m_compiled(cfg->nz / 2, fp32))
            return fail(TG_ERR_SHAPE,
                        "grid %dx%dx%d unsupported: Nx, Ny in [8, 1024] and Nz/2 in [8, 512], each 2^a (FP32), or "
                        "2^a, 3*2^a, 5*2^a or 15*2^a and a multiple of 8 (FP64)",
                        cfg->nx, cfg->ny, cfg->nz);
    }
    if (cfg->dealias < 0 || cfg->dealias > 2) return fail(TG_ERR_ARG, "unknown dealias mode %d", cfg->dealias);
    for (int i = 0; i < 3; i++)
        if (!(cfg->kfactor[i] > 0)) return fail(TG_ERR_ARG, "kfactor[%d] must be positive", i);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(TG_ERR_HIP, "no HIP device available");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(TG_ERR_ARG, "device %d out of range", cfg->device);
    tg_plan *p = new (std::nothrow) tg_plan();
    if (!p) return fail(TG_ERR_OOM, "host allocation failed");
    p->cfg = *cfg;
    p->g.nx = cfg->nx;
    p->g.ny = cfg->ny;
    p->g.nz = cfg->nz;
    p->g.nzc = cfg->nz / 2 + 1;
    for (int i = 0; i < 3; i++) p->g.kf[i] = cfg->kfactor[i];
    p->g.dealias = cfg->dealias;
    // ChFF runs its x transforms on the cosine passes of the SFF basis
    p->g.basis = cfg->basis == TG_BASIS_FFF ? 0 : 1;
    p->prec = cfg->precision;
    p->elem = p->prec == TG_PREC_FP32 ? 8 : 16;
    {
        const int per_line = 128 / (int)p->elem;  // elements per 128-byte line
        p->g.nzp = (p->g.nzc + per_line - 1) / per_line * per_line;
    }
    if (const char *e = getenv("TG_X_TILE_COLUMNS")) p->x_tile_columns = atoi(e);
    p->nc = (size_t)cfg->nx * cfg->ny * (cfg->nz / 2 + 1);
    p->ncw = (size_t)cfg->nx * cfg->ny * p->g.nzp;
    int rc = set_device(p);
    if (rc) { free_plan(p); return rc; }
    if (p->prec == TG_PREC_FP32) {
        rc = upload_twiddles<float>(p, &p->twx, cfg->nx, cfg->nx);
        if (!rc) rc = upload_twiddles<float>(p, &p->twy, cfg->ny, cfg->ny);
        if (!rc) rc = upload_twiddles<float>(p, &p->twm, cfg->nz / 2, cfg->nz / 2);
        if (!rc) rc = upload_twiddles<float>(p, &p->twn, cfg->nz, cfg->nz / 2 + 1);
        if (!rc && p->g.basis == 1) rc = upload_twiddles<float>(p, &p->twq, 4 * cfg->nx, cfg->nx);
    } else {
        rc = upload_twiddles<double>(p, &p->twx, cfg->nx, cfg->nx);
        if (!rc) rc = upload_twiddles<double>(p, &p->twy, cfg->ny, cfg->ny);
        if (!rc) rc = upload_twiddles<double>(p, &p->twm, cfg->nz / 2, cfg->nz / 2);
        if (!rc) rc = upload_twiddles<double>(p, &p->twn, cfg->nz, cfg->nz / 2 + 1);
        if (!rc && p->g.basis == 1) rc = upload_twiddles<double>(p, &p->twq, 4 * cfg->nx, cfg->nx);
    }
    for (int i = 0; i < 6 && !rc; i++) rc = dmalloc(p, &p->U[i], p->ncw * p->elem);
    for (int i = 0; i < 9 && !rc; i++) rc = dmalloc(p, &p->P[i], p->ncw * p->elem);
    // the pad columns kz >= Nz/2+1 of the work fields are never written; the x
    // passes load them with their tiles (and discard the results): keep them zero
    for (int i = 0; i < 6 && !rc; i++)
        if (hipMemset(p->U[i], 0, p->ncw * p->elem) != hipSuccess) rc = fail(TG_ERR_HIP, "hipMemset failed");
    for (int i = 0; i < 9 && !rc; i++)
        if (hipMemset(p->P[i], 0, p->ncw * p->elem) != hipSuccess) rc = fail(TG_ERR_HIP, "hipMemset failed");
    // accumulator: FP64, padded work layout (the RK state accumulation is FP64
    // in both precisions; also holds giver outputs in the boundary layout)
    for (int i = 0; i < 3 && !rc; i++) rc = dmalloc(p, (void **)&p->acc[i], p->ncw * 16);
    p->red_blocks = (int)std::min<size_t>(2048, (p->nc + 1023) / 1024);
    if (p->red_blocks < 1) p->red_blocks = 1;
    p->partial_bytes = (size_t)p->red_blocks * 2 * 1100 * sizeof(double);
    if (!rc) rc = dmalloc(p, (void **)&p->partial, p->partial_bytes);
    if (!rc) rc = dmalloc(p, (void **)&p->small, 65536);
    if (!rc) rc = dmalloc(p, (void **)&p->nonfinite_dev, 256);
    if (!rc && hipHostMalloc((void **)&p->nonfinite_host, sizeof(int), hipHostMallocDefault) != hipSuccess) {
        p->nonfinite_host = nullptr;
        rc = fail(TG_ERR_OOM, "hipHostMalloc failed");
    }
    if (rc) { free_plan(p); return rc; }
    *plan = p;
    return TG_OK;
}

int tg_plan_destroy(tg_plan *plan)
{
    if (!plan) return TG_OK;
    (void)hipSetDevice(plan->cfg.device);
    (void)hipDeviceSynchronize();
    free_plan(plan);
    return TG_OK;
}

int tg_set_stream(tg_plan *plan, void *hip_stream)
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    plan->stream = (hipStream_t)hip_stream;
    return TG_OK;
}

int tg_synchronize(tg_plan *plan)
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    HIPCHK(hipStreamSynchronize(plan->stream));
    return TG_OK;
}

int tg_get_shapes(const tg_plan *plan, int64_t cshape[3], int64_t rshape[3])
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    if (cshape) { cshape[0] = plan->g.nx; cshape[1] = plan->g.ny; cshape[2] = plan->g.nzc; }
    if (rshape) { rshape[0] = plan->g.nx; rshape[1] = plan->g.ny; rshape[2] = plan->g.nz + 2; }
    return TG_OK;
}

int64_t tg_plan_device_bytes(const tg_plan *plan) { return plan ? plan->dev_bytes : 0; }

static int ensure_chff(tg_plan *p);

static int check_sincos(tg_plan *p, int code)
{
    if (p->cfg.basis == TG_BASIS_FFF && code != TG_SC_FOURIER) return fail(TG_ERR_ARG, "FFF basis takes sincos_code 0");
    if (p->cfg.basis == TG_BASIS_SFF && code != TG_SC_SIN && code != TG_SC_COS)
        return fail(TG_ERR_ARG, "SFF basis takes sincos_code TG_SC_SIN or TG_SC_COS");
    if (p->cfg.basis == TG_BASIS_CHFF) {
        if (code != TG_SC_CHEB) return fail(TG_ERR_ARG, "ChFF basis takes sincos_code TG_SC_CHEB");
        return ensure_chff(p);
    }
    return TG_OK;
}

static int need_basis(const tg_plan *p, int basis, const char *what)
{
    static const char *names[3] = {"TG_BASIS_FFF", "TG_BASIS_SFF", "TG_BASIS_CHFF"};
    if (p->cfg.basis != basis) return fail(TG_ERR_ARG, "%s needs a plan with basis %s", what, names[basis]);
    return TG_OK;
}

static int not_chff(const tg_plan *p, const char *what)
{
    if (p->cfg.basis == TG_BASIS_CHFF) return fail(TG_ERR_ARG, "%s is not defined on ChFF plans", what);
    return TG_OK;
}

static int check_rbc(const tg_rbc *r)
{
    if (!r) return fail(TG_ERR_ARG, "null tg_rbc");
    if (!(r->Ra > 0) || !(r->Pr > 0) || !(r->dt > 0)) return fail(TG_ERR_ARG, "need Ra > 0, Pr > 0, dt > 0");
    return TG_OK;
}

int tg_forward(tg_plan *plan, int sincos_code, const double *real_dev, double *spec_dev)
{
    if (!plan || !real_dev || !spec_dev) return fail(TG_ERR_ARG, "null argument");
    int rc = check_sincos(plan, sincos_code);
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    plan->unext_valid = false;
    return DISPATCH_PREC(plan, run_forward)(plan, real_dev, spec_dev, sincos_code);
}

int tg_inverse(tg_plan *plan, int sincos_code, const double *spec_dev, double *real_dev)
{
    if (!plan || !real_dev || !spec_dev) return fail(TG_ERR_ARG, "null argument");
    int rc = check_sincos(plan, sincos_code);
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    plan->unext_valid = false;
    return DISPATCH_PREC(plan, run_inverse)(plan, spec_dev, real_dev, sincos_code);
}

int tg_get_dt(tg_plan *plan, double *const V_dev[3], double courant, double dt_fixed, double *dt, double umax[3])
{
    if (!plan || !dt) return fail(TG_ERR_ARG, "null argument");
    if (int rb_ = not_chff(plan, "tg_get_dt")) return rb_;
    if (!(courant > 0) || !(dt_fixed > 0)) return fail(TG_ERR_ARG, "need courant > 0 and dt_fixed > 0");
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    tg_plan *p = plan;
    if (!p->rfield) rc = dmalloc(p, (void **)&p->rfield, (size_t)p->g.nx * p->g.ny * (p->g.nz + 2) * 8);
    if (!rc && !p->umax_dev) rc = dmalloc(p, (void **)&p->umax_dev, 256);
    if (rc) return rc;
    // Compute_nlin dealiases U.cvf before its Inverse_transform (compute_nlin.cc:60-64)
    if (p->g.dealias) HIPCHK(launch_dealias(p->g, V_dev, 3, p->stream));
    p->unext_valid = false;
    const bool sff = p->g.basis == 1;
    for (int c = 0; c < 3; c++) {
        const int sc = sff ? (c == 0 ? TG_SC_SIN : TG_SC_COS) : 0;
        rc = DISPATCH_PREC(p, run_inverse)(p, V_dev[c], p->rfield, sc);
        if (rc) return rc;
        HIPCHK(launch_max_abs_real(p->g, p->rfield, p->umax_dev + c, p->stream));
    }
    double um[3];
    HIPCHK(hipMemcpyAsync(um, p->umax_dev, sizeof(um), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    if (umax)
        for (int c = 0; c < 3; c++) umax[c] = um[c];
    for (int c = 0; c < 3; c++)
        if (!std::isfinite(um[c])) return fail(TG_ERR_NAN, "tg_get_dt: the real-space velocity is not finite");
    // Delta_x_i = L_i / N_i, L_i = 2 pi / kfactor_i (FFF_pencil.cc:64-72; SFF L_x = pi / kfactor_x,
    // SFF_pencil.cc:77-86)
    const double L[3] = {(sff ? M_PI : 2 * M_PI) / p->g.kf[0], 2 * M_PI / p->g.kf[1], 2 * M_PI / p->g.kf[2]};
    const double dx = L[0] / p->g.nx, dy = L[1] / p->g.ny, dz = L[2] / p->g.nz;
    double sum = um[0] / dx + um[2] / dz;
    if (p->g.ny > 1) sum += um[1] / dy;
    const double d = courant / sum;
    *dt = d < dt_fixed ? d : dt_fixed;
    return TG_OK;
}

int tg_dealias(tg_plan *plan, double *const V_dev[], int ncomp)
{
    if (!plan || ncomp < 1 || ncomp > 3) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = not_chff(plan, "tg_dealias")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, ncomp, "tg_dealias");
    if (rc) return rc;
    rc = set_device(plan);
    if (rc) return rc;
    if (plan->g.dealias) HIPCHK(launch_dealias(plan->g, V_dev, ncomp, plan->stream));
    return TG_OK;
}

int tg_strong_reality(tg_plan *plan, double *const V_dev[], int ncomp)
{
    if (!plan || ncomp < 1 || ncomp > 3) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = not_chff(plan, "tg_strong_reality")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, ncomp, "tg_strong_reality");
    if (rc) return rc;
    rc = set_device(plan);
    if (rc) return rc;
    plan->unext_valid = false;
    if (plan->g.basis == 1) HIPCHK(launch_sff_reality(plan->g, V_dev, ncomp, 1, plan->stream));
    else HIPCHK(launch_reality(plan->g, V_dev, ncomp, plan->stream));
    return TG_OK;
}

int tg_weak_reality(tg_plan *plan, double *const V_dev[], int ncomp)
{
    if (!plan || ncomp < 1 || ncomp > 4) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = not_chff(plan, "tg_weak_reality")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, ncomp, "tg_weak_reality");
    if (rc) return rc;
    rc = set_device(plan);
    if (rc) return rc;
    plan->unext_valid = false;
    // the kz = Nz/2 plane is zeroed in both bases
    HIPCHK(launch_sff_reality(plan->g, V_dev, ncomp, 2, plan->stream));
    return TG_OK;
}

int tg_compute_nlin_rbc(tg_plan *plan, double *const V_dev[3], double *T_dev, double *const nlin_dev[3],
                        double *nlinT_dev)
{
    if (!plan || !T_dev || !nlinT_dev) return fail(TG_ERR_ARG, "null argument");
    int rc = need_basis(plan, 1, "tg_compute_nlin_rbc");
    if (!rc) rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)nlin_dev, 3, "nlin");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_nlin_rbc)(plan, V_dev, T_dev, nlin_dev, nlinT_dev, nullptr);
}

int tg_compute_rhs_rbc(tg_plan *plan, double *const V_dev[3], double *T_dev, const tg_rbc *rbc,
                       double *const rhs_dev[3], double *rhsT_dev)
{
    if (!plan || !T_dev || !rhsT_dev) return fail(TG_ERR_ARG, "null argument");
    int rc = need_basis(plan, 1, "tg_compute_rhs_rbc");
    if (!rc) rc = check_rbc(rbc);
    if (!rc) rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)rhs_dev, 3, "rhs");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_nlin_rbc)(plan, V_dev, T_dev, rhs_dev, rhsT_dev, rbc);
}

int tg_rk4_step_rbc(tg_plan *plan, double *const V_dev[3], double *T_dev, const tg_rbc *rbc, int nsteps)
{
    if (!plan || !T_dev) return fail(TG_ERR_ARG, "null argument");
    if (nsteps < 0) return fail(TG_ERR_ARG, "nsteps must be >= 0");
    int rc = need_basis(plan, 1, "tg_rk4_step_rbc");
    if (!rc) rc = check_rbc(rbc);
    if (!rc) rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_rk4_rbc)(plan, V_dev, T_dev, rbc, nsteps);
}

int tg_nusselt(tg_plan *plan, const double *V1_dev, const double *T_dev, const tg_rbc *rbc, double *nu)
{
    if (!plan || !V1_dev || !T_dev || !nu) return fail(TG_ERR_ARG, "null argument");
    int rc = need_basis(plan, 1, "tg_nusselt");
    if (!rc) rc = check_rbc(rbc);
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    const double *A[1] = {V1_dev}, *B[1] = {T_dev};
    double e = 0.0;
    rc = reduce_to_host(plan, RED_CROSS, A, B, 1, 0, nullptr, 1, 1, 0, 0, 0, &e);
    if (rc) return rc;
    // RED_CROSS carries the SFF weights w = (lx?2:1)(lz?2:1)/2 of
    // SFF_pencil_energy.cc:117-133
    *nu = 1.0 + 2.0 * std::sqrt(rbc->Ra * rbc->Pr) * e;
    return TG_OK;
}

int tg_compute_nlin(tg_plan *plan, double *const V_dev[3], double *const nlin_dev[3])
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    if (int rb_ = need_basis(plan, 0, "tg_compute_nlin")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)nlin_dev, 3, "nlin");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_nlin)(plan, V_dev, nlin_dev, nullptr, 0);
}

int tg_compute_rhs(tg_plan *plan, double *const V_dev[3], const tg_fluid *fluid, double *const rhs_dev[3])
{
    if (!plan || !fluid) return fail(TG_ERR_ARG, "null argument");
    if (int rb_ = need_basis(plan, 0, "tg_compute_rhs")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)rhs_dev, 3, "rhs");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_nlin)(plan, V_dev, rhs_dev, fluid, 1);
}

int tg_compute_nlin_giver(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], double *const nlin_dev[3])
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    if (int rb_ = need_basis(plan, 0, "tg_compute_nlin_giver")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)W_dev, 3, "W");
    if (!rc) rc = check_ptrs((const void *const *)nlin_dev, 3, "nlin");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    if (plan->g.dealias) {
        HIPCHK(launch_dealias(plan->g, V_dev, 3, plan->stream));
        HIPCHK(launch_dealias(plan->g, W_dev, 3, plan->stream));
    }
    return DISPATCH_PREC(plan, run_giver_nlin)(plan, V_dev, W_dev, 0, 0.0, 0.0, nlin_dev, false, 0.0, 0.0, 3);
}

int tg_rk4_step(tg_plan *plan, double *const V_dev[3], const tg_fluid *fluid, int nsteps, int flags)
{
    if (!plan || !fluid) return fail(TG_ERR_ARG, "null argument");
    if (int rb_ = need_basis(plan, 0, "tg_rk4_step")) return rb_;
    if (nsteps < 0) return fail(TG_ERR_ARG, "nsteps must be >= 0");
    if (!(fluid->dt > 0) || !(fluid->nu >= 0)) return fail(TG_ERR_ARG, "need dt > 0 and nu >= 0");
    if (int rh_ = check_hyper(fluid, true)) return rh_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_rk4)(plan, V_dev, fluid, nsteps, flags);
}

int tg_energy(tg_plan *plan, const double *A_dev, double *energy)
{
    if (!plan || !A_dev || !energy) return fail(TG_ERR_ARG, "null argument");
    if (int rb_ = not_chff(plan, "tg_energy")) return rb_;
    int rc = set_device(plan);
    if (rc) return rc;
    const double *A[1] = {A_dev};
    return reduce_to_host(plan, RED_ENERGY, A, nullptr, 1, 0, nullptr, 1, 1, 0, 0, 0, energy);
}

int tg_Sn(tg_plan *plan, const double *A_dev, int n, double *Sn)
{
    if (!plan || !A_dev || !Sn || n < 0) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_Sn")) return rb_;
    int rc = set_device(plan);
    if (rc) return rc;
    const double *A[1] = {A_dev};
    return reduce_to_host(plan, RED_SN, A, nullptr, 1, n, nullptr, 1, 1, min_radius_outside(plan->g), 0, 0, Sn);
}

int tg_shell_spectrum(tg_plan *plan, const double *A_dev, int n, double *Sk, int nbins)
{
    if (!plan || !A_dev || !Sk || n < 0) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_shell_spectrum")) return rb_;
    const int kmax = min_radius_outside(plan->g);
    if (nbins < kmax + 1) return fail(TG_ERR_ARG, "nbins must be >= %d", kmax + 1);
    int rc = set_device(plan);
    if (rc) return rc;
    std::vector<double> raw(2 * (size_t)(kmax + 1));
    const double *A[1] = {A_dev};
    rc = reduce_to_host(plan, RED_SPECTRUM, A, nullptr, 1, n, nullptr, kmax + 1, 2, kmax, 0, 0, raw.data());
    if (rc) return rc;
    const double kf = plan->g.kf[0] * plan->g.kf[1] * plan->g.kf[2];
    for (int i = 0; i < nbins; i++) Sk[i] = 0.0;
    for (int i = 0; i <= kmax; i++) {
        const double cnt = raw[(size_t)(kmax + 1) + i];
        Sk[i] = cnt <= 0 ? 0.0 : raw[i] * (4 * M_PI) / (cnt * kf);
    }
    return TG_OK;
}

int tg_flux(tg_plan *plan, double *const V_dev[3], const double *radii, int nspheres, double *flux)
{
    if (!plan || !radii || !flux || nspheres < 1) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_flux")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    // Compute_nlin(U, Giver) dealiases U in place (compute_nlin_ET_VF.cc:74-77)
    if (plan->g.dealias) HIPCHK(launch_dealias(plan->g, V_dev, 3, plan->stream));
    flux[0] = 0.0;
    for (int s = 1; s <= nspheres; s++) {
        rc = DISPATCH_PREC(plan, run_giver_nlin)(plan, V_dev, V_dev, 8, 0.0, radii[s], plan->acc, s > 1, 0.0, 0.0, 3);
        if (rc) return rc;
        double v = 0.0;
        const double *A[3] = {plan->acc[0], plan->acc[1], plan->acc[2]};
        const double *B[3] = {V_dev[0], V_dev[1], V_dev[2]};
        rc = reduce_to_host(plan, RED_SHELL_SINGLE, A, B, 3, 0, nullptr, 1, 1, 0, radii[s], INFINITY, &v);
        if (rc) return rc;
        flux[s] = -v;
    }
    return TG_OK;
}

int tg_shell_to_shell(tg_plan *plan, double *const V_dev[3], const double *radii, int nshells, double *S)
{
    if (!plan || !radii || !S || nshells < 1) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_shell_to_shell")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    const int nr = nshells + 1;
    if (nr > 1000) return fail(TG_ERR_ARG, "too many shells");
    if (plan->g.dealias) HIPCHK(launch_dealias(plan->g, V_dev, 3, plan->stream));
    for (int i = 0; i < nr * nr; i++) S[i] = 0.0;
    std::vector<double> row(nr);
    const int kin = max_radius_inside(plan->g);
    for (int m = 1; m < nshells; m++) {
        rc = DISPATCH_PREC(plan, run_giver_nlin)(plan, V_dev, V_dev, 8, radii[m - 1], radii[m], plan->acc,
                                                 m > 1, 0.0, 0.0, 3);
        if (rc) return rc;
        const double *A[3] = {plan->acc[0], plan->acc[1], plan->acc[2]};
        const double *B[3] = {V_dev[0], V_dev[1], V_dev[2]};
        rc = reduce_to_host(plan, RED_SHELL_ALL, A, B, 3, 0, radii, nr, 1, kin, 0, 0, row.data());
        if (rc) return rc;
        for (int n = 0; n < nr; n++) S[(size_t)m * nr + n] = -row[n];
    }
    return TG_OK;
}

int tg_energy_transfer(tg_plan *plan, double *const V_dev[3], const double *radii, int n, double *flux, double *S)
{
    if (!plan || !radii || !S || !flux || n < 2) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_energy_transfer")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    const int nr = n + 1;
    if (nr > 1000) return fail(TG_ERR_ARG, "too many shells");
    // the flux is assembled from the shell-to-shell givers 1..n-1: every mode
    // must lie in one of them or in the receiver-only last shell, i.e. the
    // first sphere is the origin and the last one encloses every mode
    // (Min_radius_outside); radii non-decreasing (the default radii repeat
    // a value when their log spacing is finer than the grid, leaving empty
    // shells)
    if (!(radii[0] == 0.0)) return fail(TG_ERR_ARG, "tg_energy_transfer needs radii[0] == 0");
    for (int i = 1; i <= n; i++)
        if (!(radii[i] >= radii[i - 1])) return fail(TG_ERR_ARG, "radii must not decrease");
    if (!(radii[n] >= (double)min_radius_outside(plan->g)))
        return fail(TG_ERR_ARG, "tg_energy_transfer needs radii[n] >= %d (a last sphere enclosing every mode)",
                    min_radius_outside(plan->g));
    if (plan->g.dealias) HIPCHK(launch_dealias(plan->g, V_dev, 3, plan->stream));
    for (int i = 0; i < nr * nr; i++) S[i] = 0.0;
    // T[m][k]: transfer from giver shell m into receiver shell k, all modes
    // (the flux's receiver includes the modes beyond the largest inscribed
    // sphere, which the shell-to-shell matrix leaves out)
    std::vector<double> row(nr), T((size_t)nr * nr, 0.0);
    const int kin = max_radius_inside(plan->g);
    for (int m = 1; m < n; m++) {
        rc = DISPATCH_PREC(plan, run_giver_nlin)(plan, V_dev, V_dev, 8, radii[m - 1], radii[m], plan->acc,
                                                 m > 1, 0.0, 0.0, 3);
        if (rc) return rc;
        const double *A[3] = {plan->acc[0], plan->acc[1], plan->acc[2]};
        const double *B[3] = {V_dev[0], V_dev[1], V_dev[2]};
        rc = reduce_to_host(plan, RED_SHELL_ALL, A, B, 3, 0, radii, nr, 1, kin, 0, 0, row.data());
        if (rc) return rc;
        for (int k = 0; k < nr; k++) S[(size_t)m * nr + k] = -row[k];
        rc = reduce_to_host(plan, RED_SHELL_ALL_NOCUT, A, B, 3, 0, radii, nr, 1, kin, 0, 0, row.data());
        if (rc) return rc;
        for (int k = 0; k < nr; k++) T[(size_t)m * nr + k] = -row[k];
    }
    // flux through sphere s: givers inside (shells 1..s), receivers outside
    flux[0] = 0.0;
    for (int s = 1; s <= n; s++) {
        double f = 0.0;
        for (int m = 1; m <= std::min(s, n - 1); m++)
            for (int k = s + 1; k < nr; k++) f += T[(size_t)m * nr + k];
        flux[s] = f;
    }
    return TG_OK;
}

int tg_ring_spectrum(tg_plan *plan, const double *const A_dev[], int ncomp, int n, int dirn,
                     const double *sector_angles, int nsectors, double *out)
{
    if (!plan || !A_dev || !out || n < 0 || (ncomp != 1 && ncomp != 3)) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_ring_spectrum")) return rb_;
    if (dirn < 1 || dirn > 3) return fail(TG_ERR_ARG, "anisotropy direction must be 1, 2 or 3");
    int rc = check_ptrs((const void *const *)A_dev, ncomp, "A");
    if (!rc) rc = check_edges(sector_angles, nsectors, "sector angles");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    const int kmax = max_radius_inside(plan->g);
    return reduce_aniso_to_host(plan, RED_RING_SPEC, A_dev, nullptr, ncomp, n, ncomp, nullptr, 0, sector_angles,
                                nsectors, dirn, (kmax + 1) * (nsectors + 1), kmax, out);
}

int tg_cylindrical_ring_spectrum(tg_plan *plan, const double *const V_dev[3], int n, int dirn, const double *kpll,
                                 int nslabs, double *S1, double *S2)
{
    if (!plan || !V_dev || !S1 || !S2 || n < 0) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_cylindrical_ring_spectrum")) return rb_;
    if (dirn < 1 || dirn > 3) return fail(TG_ERR_ARG, "anisotropy direction must be 1, 2 or 3");
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_edges(kpll, nslabs, "Kpll edges");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    const int kmax = aniso_max_krho_inside(plan->g, dirn);
    const int nb = (kmax + 1) * (nslabs + 1);
    std::vector<double> tmp(2 * (size_t)nb);
    rc = reduce_aniso_to_host(plan, RED_CYL_SPEC, V_dev, nullptr, 3, n, 2, nullptr, 0, kpll, nslabs, dirn, nb, kmax,
                              tmp.data());
    if (rc) return rc;
    std::copy(tmp.begin(), tmp.begin() + nb, S1);
    std::copy(tmp.begin() + nb, tmp.end(), S2);
    return TG_OK;
}

// Ring-to-ring (mode 16) and cylindrical-ring-to-ring (mode 32) transfers:
// one giver evaluation per (shell, sub-bin) giver ring, binned receivers.
static int aniso_transfer(tg_plan *plan, double *const V_dev[3], int dirn, const double *radii, int nshells,
                          const double *sub, int nsub, double *out, bool cyl, const char *what)
{
    if (!plan || !V_dev || !out) return fail(TG_ERR_ARG, "null argument");
    if (int rb_ = need_basis(plan, 0, what)) return rb_;
    if (dirn < 1 || dirn > 3) return fail(TG_ERR_ARG, "anisotropy direction must be 1, 2 or 3");
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_edges(radii, nshells, "radii");
    if (!rc) rc = check_edges(sub, nsub, cyl ? "Kpll edges" : "sector angles");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    const size_t rb = (size_t)(nshells + 1) * (nsub + 1);
    if (rb > 4096) return fail(TG_ERR_ARG, "too many rings (%zu)", rb);
    // Compute_nlin(U, Giver) dealiases U in place (compute_nlin_ET_VF.cc:74-77)
    if (plan->g.dealias) HIPCHK(launch_dealias(plan->g, V_dev, 3, plan->stream));
    std::fill(out, out + rb * rb, 0.0);
    const int kcut = cyl ? aniso_max_krho_inside(plan->g, dirn) : max_radius_inside(plan->g);
    std::vector<double> row(rb);
    const double *B[3] = {V_dev[0], V_dev[1], V_dev[2]};
    const double *A[3] = {plan->acc[0], plan->acc[1], plan->acc[2]};
    bool first = true;
    // the last shell (outer radius infinite) gives nothing (compute_ring_tr.cc:64)
    for (int i = 1; i < nshells; i++)
        for (int j = 1; j <= nsub; j++) {
            rc = DISPATCH_PREC(plan, run_giver_nlin)(plan, V_dev, V_dev, cyl ? 32 : 16, radii[i - 1], radii[i],
                                                     plan->acc, !first, sub[j - 1], sub[j], dirn);
            if (rc) return rc;
            first = false;
            rc = reduce_aniso_to_host(plan, cyl ? RED_CYL_MULT : RED_RING_MULT, A, B, 3, 0, 1, radii, nshells, sub,
                                      nsub, dirn, (int)rb, kcut, row.data());
            if (rc) return rc;
            double *o = out + ((size_t)i * (nsub + 1) + j) * rb;
            for (size_t k = 0; k < rb; k++) o[k] = -row[k];
        }
    return TG_OK;
}

int tg_ring_to_ring(tg_plan *plan, double *const V_dev[3], int dirn, const double *radii, int nshells,
                    const double *sector_angles, int nsectors, double *out)
{
    return aniso_transfer(plan, V_dev, dirn, radii, nshells, sector_angles, nsectors, out, false, "tg_ring_to_ring");
}

int tg_cylindrical_ring_to_ring(tg_plan *plan, double *const V_dev[3], int dirn, const double *radii, int nshells,
                                const double *kpll, int nslabs, double *out)
{
    return aniso_transfer(plan, V_dev, dirn, radii, nshells, kpll, nslabs, out, true, "tg_cylindrical_ring_to_ring");
}

int tg_compute_nlin_mhd(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], double *const nlinV_dev[3],
                        double *const nlinW_dev[3])
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    if (int rb_ = need_basis(plan, 0, "tg_compute_nlin_mhd")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)W_dev, 3, "W");
    if (!rc) rc = check_ptrs((const void *const *)nlinV_dev, 3, "nlinV");
    if (!rc) rc = check_ptrs((const void *const *)nlinW_dev, 3, "nlinW");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_nlin_mhd)(plan, V_dev, W_dev, nlinV_dev, nlinW_dev);
}

int tg_rk4_step_mhd(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], const tg_fluid *fluid, double eta,
                    int nsteps)
{
    if (!plan || !fluid) return fail(TG_ERR_ARG, "null argument");
    if (int rb_ = need_basis(plan, 0, "tg_rk4_step_mhd")) return rb_;
    if (nsteps < 0) return fail(TG_ERR_ARG, "nsteps must be >= 0");
    if (!(fluid->dt > 0) || !(fluid->nu >= 0) || !(eta >= 0)) return fail(TG_ERR_ARG, "need dt > 0, nu, eta >= 0");
    if (int rh_ = check_hyper(fluid, false)) return rh_;
    if (fluid->force_kind != 0) return fail(TG_ERR_ARG, "forcing is not supported for MHD");
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)W_dev, 3, "W");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_rk4_mhd)(plan, V_dev, W_dev, fluid, eta, nsteps);
}

int tg_cross_helicity(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], double *hc)
{
    if (!plan || !hc) return fail(TG_ERR_ARG, "null argument");
    if (int rb_ = need_basis(plan, 0, "tg_cross_helicity")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)W_dev, 3, "W");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    double t = 0.0;
    for (int c = 0; c < 3; c++) {
        double v = 0.0;
        const double *A[1] = {V_dev[c]};
        const double *B[1] = {W_dev[c]};
        rc = reduce_to_host(plan, RED_CROSS, A, B, 1, 0, nullptr, 1, 1, 0, 0, 0, &v);
        if (rc) return rc;
        t += v;
    }
    *hc = t;
    return TG_OK;
}

int tg_flux_mhd(tg_plan *plan, double *const V_dev[3], double *const W_dev[3], const double *radii, int nspheres,
                double *u2b, double *b2b)
{
    if (!plan || !radii || !u2b || !b2b || nspheres < 1) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_flux_mhd")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)W_dev, 3, "W");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    if (plan->g.dealias) {
        HIPCHK(launch_dealias(plan->g, V_dev, 3, plan->stream));
        HIPCHK(launch_dealias(plan->g, W_dev, 3, plan->stream));
    }
    const double *A[3] = {plan->acc[0], plan->acc[1], plan->acc[2]};
    const double *B[3] = {W_dev[0], W_dev[1], W_dev[2]};
    u2b[0] = b2b[0] = 0.0;
    for (int s = 1; s <= nspheres; s++) {
        // W.nlin = FT[d_j(b_j u<_i)]; u< -> b> = sum_{K > r} 2m Re(N.b*)
        rc = DISPATCH_PREC(plan, run_giver_nlin)(plan, W_dev, V_dev, 8, 0.0, radii[s], plan->acc, false, 0.0, 0.0, 3);
        if (rc) return rc;
        rc = reduce_to_host(plan, RED_SHELL_SINGLE, A, B, 3, 0, nullptr, 1, 1, 0, radii[s], INFINITY, &u2b[s]);
        if (rc) return rc;
    }
    for (int s = 1; s <= nspheres; s++) {
        // U.nlin = FT[d_j(u_j b<_i)]; b< -> b> = -sum_{K > r} 2m Re(N.b*)
        rc = DISPATCH_PREC(plan, run_giver_nlin)(plan, V_dev, W_dev, 8, 0.0, radii[s], plan->acc, false, 0.0, 0.0, 3);
        if (rc) return rc;
        double v = 0.0;
        rc = reduce_to_host(plan, RED_SHELL_SINGLE, A, B, 3, 0, nullptr, 1, 1, 0, radii[s], INFINITY, &v);
        if (rc) return rc;
        b2b[s] = -v;
    }
    return TG_OK;
}

int tg_rhs_vjp(tg_plan *plan, double *const V_dev[3], double *const lam_dev[3], double *const out_dev[3])
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    if (int rb_ = need_basis(plan, 0, "tg_rhs_vjp")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)lam_dev, 3, "lam");
    if (!rc) rc = check_ptrs((const void *const *)out_dev, 3, "out");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_rhs_vjp)(plan, V_dev, lam_dev, out_dev);
}

int tg_adjoint_forward(tg_plan *plan, double *const V_dev[3], const tg_fluid *fluid, int nsteps, int checkpoint_every)
{
    if (!plan || !fluid) return fail(TG_ERR_ARG, "null argument");
    if (int rb_ = need_basis(plan, 0, "tg_adjoint_forward")) return rb_;
    if (nsteps < 1 || checkpoint_every < 1) return fail(TG_ERR_ARG, "need nsteps >= 1 and checkpoint_every >= 1");
    if (!(fluid->dt > 0) || !(fluid->nu >= 0)) return fail(TG_ERR_ARG, "need dt > 0 and nu >= 0");
    if (int rh_ = check_hyper(fluid, true)) return rh_;
    if (fluid->force_kind != 0) return fail(TG_ERR_ARG, "the adjoint supports force_kind 0 only");
    if (plan->cfg.strong_reality) return fail(TG_ERR_ARG, "the adjoint needs a plan with strong_reality 0");
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_adjoint_forward)(plan, V_dev, fluid, nsteps, checkpoint_every);
}

int tg_adjoint_backward(tg_plan *plan, double *const lam_dev[3])
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    if (!plan->adj.recorded) return fail(TG_ERR_STATE, "no tg_adjoint_forward sweep recorded on this plan");
    int rc = check_ptrs((const void *const *)lam_dev, 3, "lam");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    return DISPATCH_PREC(plan, run_adjoint_backward)(plan, lam_dev);
}

int tg_adjoint_free(tg_plan *plan)
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    int rc = set_device(plan);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(plan->stream));
    adj_free(plan);
    return TG_OK;
}

int tg_profile_enable(tg_plan *plan, int on)
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    int rc = prof_collect(plan);
    if (rc) return rc;
    plan->prof_on = on != 0;
    return TG_OK;
}

int tg_profile_reset(tg_plan *plan)
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    int rc = prof_collect(plan);
    if (rc) return rc;
    for (int k = 0; k < TG_KERNEL_COUNT; k++) { plan->prof_ms[k] = 0; plan->prof_n[k] = 0; }
    return TG_OK;
}

int tg_profile_read(tg_plan *plan, int kernel, double *total_ms, int64_t *launches)
{
    if (!plan || kernel < 0 || kernel >= TG_KERNEL_COUNT) return fail(TG_ERR_ARG, "bad argument");
    int rc = prof_collect(plan);
    if (rc) return rc;
    if (total_ms) *total_ms = plan->prof_ms[kernel];
    if (launches) *launches = plan->prof_n[kernel];
    return TG_OK;
}

const char *tg_kernel_name(int kernel)
{
    static const char *names[TG_KERNEL_COUNT] = {"x_inv", "y_fft", "z_products", "x_nlin_rk", "x_nlin_giver",
                                                 "reality", "z_r2c", "z_c2r", "x_fwd", "reduce", "x_inv_sym",
                                                 "z_adj", "x_fwd_adj", "x_sc", "z_rbc", "x_rbc"};
    return (kernel >= 0 && kernel < TG_KERNEL_COUNT) ? names[kernel] : "unknown";
}

int tg_init_field(tg_plan *plan, int kind, double k0, double kmin, double kmax, double energy, uint64_t seed,
                  double *const V_dev[3])
{
    if (!plan || (kind != 0 && kind != 1) || !(energy >= 0)) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, 0, "tg_init_field")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (rc) return rc;
    plan->unext_valid = false;
    HIPCHK(launch_init_field(plan->g, V_dev, kind, k0, kmin, kmax, seed, plan->stream));
    HIPCHK(launch_reality(plan->g, V_dev, 3, plan->stream));
    // energy of each component into small[1..3], target into small[0]
    double *tgt = plan->small + 2048;
    HIPCHK(hipMemcpyAsync(tgt, &energy, sizeof(double), hipMemcpyHostToDevice, plan->stream));
    for (int c = 0; c < 3; c++) {
        const double *A[1] = {V_dev[c]};
        HIPCHK(launch_reduce(plan->g, RED_ENERGY, A, nullptr, 1, 0, nullptr, 1, 1, 0, 0, 0, plan->partial,
                             plan->red_blocks, tgt + 1 + c, plan->stream));
    }
    HIPCHK(launch_scale_fields(plan->g, V_dev, tgt, plan->stream));
    HIPCHK(hipStreamSynchronize(plan->stream));
    return TG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ChFF: Chebyshev x-derivative, Helmholtz solves, nonlinear term, BDF1.
// ---------------------------------------------------------------------------
static int ensure_chff(tg_plan *p)
{
    const size_t nc16 = p->nc * 16, nr8 = (size_t)p->g.nx * p->g.ny * (p->g.nz + 2) * 8;
    const size_t ncol = (size_t)p->g.ny * p->g.nzc;
    struct { double **q; size_t bytes; } need[] = {
        {&p->ch.tmp, nc16}, {&p->ch.D, nc16}, {&p->ch.S, nc16}, {&p->ch.NL[0], nc16}, {&p->ch.NL[1], nc16},
        {&p->ch.NL[2], nc16}, {&p->ch.R[0], nr8}, {&p->ch.R[1], nr8}, {&p->ch.R[2], nr8}, {&p->ch.R[3], nr8},
        {&p->ch.X, nc16}, {&p->ch.Y, nc16}, {&p->ch.pp, p->nc * 8}, {&p->ch.pm, p->nc * 8},
        {&p->ch.vxp, p->nc * 8}, {&p->ch.vxm, p->nc * 8}, {&p->ch.M, ncol * 4 * 8}};
    for (auto &n : need)
        if (!*n.q) {
            int rc = dmalloc(p, (void **)n.q, n.bytes);
            if (rc) return rc;
        }
    return TG_OK;
}

static HelmArgs helm_args(tg_plan *p, const double *R, double *U, double lam_sup, const double vp[2],
                          const double vm[2], int E, int pressure_switch)
{
    HelmArgs h{};
    h.R = R;
    h.U = U;
    h.X = p->ch.X;
    h.Y = p->ch.Y;
    h.g = p->g;
    h.lam_sup = lam_sup;
    for (int i = 0; i < 2; i++) {
        h.vp[i] = vp[i];
        h.vm[i] = vm[i];
    }
    h.E = E;
    h.pressure_switch = pressure_switch;
    return h;
}

// N_i = sum_j d_j (u_j u_i), Chebyshev x-derivatives (no dealiasing)
static int chff_nlin(tg_plan *p, double *const V[3], double *const N[3])
{
    for (int c = 0; c < 3; c++) {
        int rc = run_inverse<double>(p, V[c], p->ch.R[c], TG_SC_CHEB);
        if (rc) return rc;
    }
    static const int pa[6] = {0, 1, 2, 0, 1, 2}, pb[6] = {0, 1, 2, 1, 2, 0};
    const size_t nr = (size_t)p->g.nx * p->g.ny * (p->g.nz + 2);
    bool assigned[3] = {false, false, false};
    for (int q = 0; q < 6; q++) {
        const int a = pa[q], b = pb[q];
        HIPCHK(launch_real_product(p->ch.R[a], p->ch.R[b], p->ch.R[3], nr, p->stream));
        int rc = run_forward<double>(p, p->ch.R[3], p->ch.S, TG_SC_CHEB);
        if (rc) return rc;
        for (int t = 0; t < (a == b ? 1 : 2); t++) {
            const int dst = t == 0 ? a : b, dir = t == 0 ? b : a;
            if (dir == 0) HIPCHK(launch_cheb_deriv(p->g, p->ch.S, p->ch.D, true, p->stream));
            HIPCHK(launch_add_deriv(p->g, N[dst], p->ch.S, p->ch.D, dir, !assigned[dst], p->stream));
            assigned[dst] = true;
        }
    }
    return TG_OK;
}

// Compute_homgeneous_soln_influence_matrix (influence_matrix.cc:74-133)
static int chff_influence(tg_plan *p, double nu, double dt)
{
    if (p->ch.infl_nu == nu && p->ch.infl_dt == dt) return TG_OK;
    double *zero = p->ch.R[3];
    HIPCHK(launch_fill_double(zero, p->nc, 0.0, p->stream));
    const double one[2] = {1.0, 0.0}, nil[2] = {0.0, 0.0};
    double *tmp = p->ch.D;  // real scratch (nc doubles)
    double *hp[2] = {p->ch.pp, p->ch.pm}, *hv[2] = {p->ch.vxp, p->ch.vxm};
    for (int s = 0; s < 2; s++) {
        HIPCHK(launch_helmholtz(helm_args(p, zero, hp[s], 0.0, s == 0 ? one : nil, s == 0 ? nil : one, 1, 1),
                                p->stream));
        HIPCHK(launch_cheb_deriv(p->g, hp[s], tmp, false, p->stream));
        HIPCHK(launch_scale_real(tmp, p->nc, -nu, p->stream));
        HIPCHK(launch_helmholtz(helm_args(p, tmp, hv[s], 1 / (dt * nu), nil, nil, 1, 0), p->stream));
    }
    HIPCHK(launch_influence(p->g, p->ch.vxp, p->ch.vxm, p->ch.M, p->stream));
    p->ch.infl_nu = nu;
    p->ch.infl_dt = dt;
    return TG_OK;
}

extern "C" {

int tg_cheb_xderiv(tg_plan *plan, const double *A_dev, double *B_dev)
{
    if (!plan || !A_dev || !B_dev || A_dev == B_dev) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, TG_BASIS_CHFF, "tg_cheb_xderiv")) return rb_;
    int rc = set_device(plan);
    if (rc) return rc;
    HIPCHK(launch_cheb_deriv(plan->g, A_dev, B_dev, true, plan->stream));
    return TG_OK;
}

int tg_helmholtz(tg_plan *plan, const double *R_dev, double *U_dev, double lambda_supplement,
                 const double value_plus1[2], const double value_minus1[2])
{
    if (!plan || !R_dev || !U_dev || !value_plus1 || !value_minus1 || R_dev == U_dev)
        return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, TG_BASIS_CHFF, "tg_helmholtz")) return rb_;
    if (plan->g.nx < 4) return fail(TG_ERR_SHAPE, "tg_helmholtz needs Nx >= 4");
    int rc = set_device(plan);
    if (!rc) rc = ensure_chff(plan);
    if (rc) return rc;
    HIPCHK(launch_helmholtz(helm_args(plan, R_dev, U_dev, lambda_supplement, value_plus1, value_minus1, 2, 0),
                            plan->stream));
    return TG_OK;
}

int tg_compute_nlin_chff(tg_plan *plan, double *const V_dev[3], double *const nlin_dev[3])
{
    if (!plan) return fail(TG_ERR_ARG, "null plan");
    if (int rb_ = need_basis(plan, TG_BASIS_CHFF, "tg_compute_nlin_chff")) return rb_;
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = check_ptrs((const void *const *)nlin_dev, 3, "nlin");
    if (!rc) rc = set_device(plan);
    if (!rc) rc = ensure_chff(plan);
    if (rc) return rc;
    return chff_nlin(plan, V_dev, nlin_dev);
}

int tg_bdf1_step(tg_plan *plan, double *const V_dev[3], double *P_dev, const tg_fluid *fluid, int nsteps)
{
    if (!plan || !P_dev || !fluid || nsteps < 0) return fail(TG_ERR_ARG, "bad argument");
    if (int rb_ = need_basis(plan, TG_BASIS_CHFF, "tg_bdf1_step")) return rb_;
    if (!(fluid->dt > 0) || !(fluid->nu > 0)) return fail(TG_ERR_ARG, "need dt > 0 and nu > 0");
    if (int rh_ = check_hyper(fluid, false)) return rh_;
    if (fluid->force_kind != 0) return fail(TG_ERR_ARG, "forcing is not supported for ChFF");
    if (plan->g.nx < 4) return fail(TG_ERR_SHAPE, "tg_bdf1_step needs Nx >= 4");
    int rc = check_ptrs((const void *const *)V_dev, 3, "V");
    if (!rc) rc = set_device(plan);
    if (!rc) rc = ensure_chff(plan);
    if (!rc) rc = chff_influence(plan, fluid->nu, fluid->dt);
    if (rc) return rc;
    const double nu = fluid->nu, dt = fluid->dt, nil[2] = {0.0, 0.0};
    tg_plan::Chff &c = plan->ch;
    for (int step = 0; step < nsteps; step++) {
        rc = chff_nlin(plan, V_dev, c.NL);
        if (rc) return rc;
        // div N, then the pressure with zero wall values (BDF1.cc:69-75)
        HIPCHK(launch_cheb_deriv(plan->g, c.NL[0], c.D, true, plan->stream));
        HIPCHK(launch_add_deriv(plan->g, c.D, c.NL[1], nullptr, 1, false, plan->stream));
        HIPCHK(launch_add_deriv(plan->g, c.D, c.NL[2], nullptr, 2, false, plan->stream));
        HIPCHK(launch_helmholtz(helm_args(plan, c.D, P_dev, 0.0, nil, nil, 2, 0), plan->stream));
        for (int k = 0; k < 3; k++) {
            if (k == 0) HIPCHK(launch_cheb_deriv(plan->g, P_dev, c.D, true, plan->stream));
            HIPCHK(launch_bdf1_rhs(plan->g, c.NL[k], V_dev[k], P_dev, c.D, k, dt, nu, plan->stream));
            HIPCHK(launch_helmholtz(helm_args(plan, c.NL[k], V_dev[k], 1 / (dt * nu), nil, nil, 2, 0), plan->stream));
            if (k == 0) {
                InflArgs ia{};
                ia.P = P_dev;
                ia.V1 = V_dev[0];
                ia.pp = c.pp;
                ia.pm = c.pm;
                ia.vxp = c.vxp;
                ia.vxm = c.vxm;
                ia.M = c.M;
                ia.nx = plan->g.nx;
                ia.ny = plan->g.ny;
                ia.nzc = plan->g.nzc;
                ia.kf = plan->g.kf[0];
                HIPCHK(launch_influence_correct(ia, plan->stream));
            }
        }
    }
    return TG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free shell spectrum: the block partials and the raw bins live in
// a static device buffer (one per GPU, held until the process ends) instead
// of plan-owned memory.  Its device address comes from hipGetSymbolAddress
// on each device's first call; the array's name is never used in host code
// as a device pointer.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kScratchDoubles = size_t(1) << 21;  // 16 MiB
constexpr int kScratchMaxDevices = 64;

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_scratch_mu;
double *g_scratch_dev[kScratchMaxDevices] = {};

// the current device's copy of `scratch` (looked up once per device)
int scratch_for_device(int dev, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= kScratchMaxDevices) return fail(TG_ERR_ARG, "device %d out of range", dev);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(TG_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_dev[dev] = (double *)ptr;
    }
    *out = g_scratch_dev[dev];
    return TG_OK;
}

// Stepanov normalisation of the raw bins, as tg_shell_spectrum does on the host
__global__ void k_spectrum_normalize(const double *raw, int kmax, double kf, double *Sk, int nbins)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbins) return;
    double v = 0.0;
    if (i <= kmax) {
        const double cnt = raw[kmax + 1 + i];
        v = cnt <= 0 ? 0.0 : raw[i] * (4 * M_PI) / (cnt * kf);
    }
    Sk[i] = v;
}

}  // namespace

extern "C" int tg_shell_spectrum_noalloc(const tg_config *cfg, const double *A_dev, int n, double *Sk_dev, int nbins,
                                         void *stream)
{
    if (!cfg || !A_dev || !Sk_dev || n < 0) return fail(TG_ERR_ARG, "bad argument");
    if (cfg->basis != TG_BASIS_FFF) return fail(TG_ERR_ARG, "tg_shell_spectrum_noalloc needs the FFF basis");
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 2) return fail(TG_ERR_SHAPE, "bad grid");
    for (int i = 0; i < 3; i++)
        if (!(cfg->kfactor[i] > 0)) return fail(TG_ERR_ARG, "kfactor[%d] must be positive", i);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev)
        return fail(TG_ERR_ARG, "device %d out of range", cfg->device);
    Grid g{};
    g.nx = cfg->nx;
    g.ny = cfg->ny;
    g.nz = cfg->nz;
    g.nzc = cfg->nz / 2 + 1;
    g.nzp = g.nzc;
    for (int i = 0; i < 3; i++) g.kf[i] = cfg->kfactor[i];
    const int kmax = min_radius_outside(g);
    if (nbins < kmax + 1) return fail(TG_ERR_ARG, "nbins must be >= %d", kmax + 1);
    // the same block count as a plan's reductions, so the sums agree bit for bit
    const size_t nc = (size_t)g.nx * g.ny * g.nzc;
    const int nblocks = (int)std::max<size_t>(1, std::min<size_t>(2048, (nc + 1023) / 1024));
    const size_t width = 2 * (size_t)(kmax + 1);
    const size_t need = (size_t)nblocks * width + width;
    if (need > kScratchDoubles)
        return fail(TG_ERR_ARG, "grid %dx%dx%d needs %zu doubles of scratch, above the static %zu (16 MiB)", g.nx,
                    g.ny, g.nz, need, kScratchDoubles);
    if (kmax + 1 > 1100) return fail(TG_ERR_ARG, "too many spectrum bins (%d)", kmax + 1);
    HIPCHK(hipSetDevice(cfg->device));
    double *buf = nullptr;
    int rc = scratch_for_device(cfg->device, &buf);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    double *partial = buf, *raw = buf + (size_t)nblocks * width;
    const double *A[1] = {A_dev};
    HIPCHK(launch_reduce(g, RED_SPECTRUM, A, nullptr, 1, n, nullptr, kmax + 1, 2, kmax, 0.0, 0.0, partial, nblocks,
                         raw, st));
    const double kf = g.kf[0] * g.kf[1] * g.kf[2];
    hipLaunchKernelGGL(k_spectrum_normalize, dim3((nbins + 255) / 256), dim3(256), 0, st, raw, kmax, kf, Sk_dev,
                       nbins);
    HIPCHK(hipGetLastError());
    return TG_OK;
}
