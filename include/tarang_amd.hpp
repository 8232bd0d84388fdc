// tarang_amd.hpp -- C++ host classes over the C ABI (tarang_gpu.h), shaped
// like the reference's basis / field / solver types so that a Tarang driver
// can switch its hot path to the GPU:
//
//   tarang::FFF_Basis                ~ FFF_PENCIL : Universal (universal.h:48-773,
//                                      FFF_pencil.cc:31-277) plus fftk's plan
//   tarang::SFF_Basis                ~ SFF_PENCIL (SFF_pencil.cc:48-102): sin/cos
//                                      along x, Fourier along y and z
//   tarang::FluidVF                  ~ FluidVF (FluidVF.h:73-180): cvf V1..V3 and
//                                      nlin1..3 on the device, dissipation_coefficient
//   tarang::FluidSF                  ~ FluidSF (FluidSF.h): csf F and nlin
//   tarang::Nlin_incompress          ~ Nlin_incompress::Compute_nlin (Nlin.h:64-112):
//                                      U; U with a giver; U, W (MHD); U, T (RBC)
//   tarang::Time_advance_incompress  ~ Time_advance_incompress (Time_advance.h:69-187):
//                                      RK4 of U; of U, W (MHD); of U, T (RBC)
//   tarang::Correlation              ~ Correlation::Get_Nusselt_no, Get_cross_helicity
//                                      (Correlation.cc:144-230)
//   tarang::EnergyTr                 ~ EnergyTr::Compute_flux / Compute_shell_tr
//                                      (EnergyTr.h:239-273), MHD flux
//   tarang::Adjoint_incompress       ~ no counterpart in the reference: the
//                                      checkpointed adjoint of the RK4 map
//
// Errors: the reference aborts (exit(1) / MPI_Abort); here every failing call
// throws tarang::Error carrying the C ABI's code and message.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "tarang_gpu.h"

namespace tarang {

struct Error : std::runtime_error {
    int code;
    Error(const char *fn, int c) : std::runtime_error(std::string(fn) + ": " + tg_last_error_string()), code(c) {}
};

inline void check(const char *fn, int rc)
{
    if (rc != TG_OK) throw Error(fn, rc);
}

// One complex spectral field (Nx, Ny, Nz/2+1) FP64 in device memory.
class DeviceField {
  public:
    DeviceField() = default;
    explicit DeviceField(size_t ncomplex) : n_(ncomplex)
    {
        if (hipMalloc(&p_, n_ * 16) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        (void)hipMemset(p_, 0, n_ * 16);
    }
    DeviceField(const DeviceField &) = delete;
    DeviceField &operator=(const DeviceField &) = delete;
    DeviceField(DeviceField &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
    DeviceField &operator=(DeviceField &&o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~DeviceField()
    {
        if (p_) (void)hipFree(p_);
    }
    double *data() const { return static_cast<double *>(p_); }
    size_t size() const { return n_; }
    void upload(const std::vector<double> &h) const { (void)hipMemcpy(p_, h.data(), n_ * 16, hipMemcpyHostToDevice); }
    std::vector<double> download() const
    {
        std::vector<double> h(2 * n_);
        (void)hipMemcpy(h.data(), p_, n_ * 16, hipMemcpyDeviceToHost);
        return h;
    }

  private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

// A plan (one basis on one device), with the Universal virtuals shared by
// the basis families.
class Basis {
  public:
    Basis(int basis, int Nx, int Ny, int Nz, std::array<double, 3> kfactor, int precision, int dealias, int reality,
          int device)
        : Nx(Nx), Ny(Ny), Nz(Nz)
    {
        tg_config c{};
        c.basis = basis;
        c.nx = Nx;
        c.ny = Ny;
        c.nz = Nz;
        for (int i = 0; i < 3; i++) c.kfactor[i] = kfactor[i];
        c.precision = precision;
        c.dealias = dealias;
        c.strong_reality = reality;
        c.device = device;
        check("tg_plan_create", tg_plan_create(&c, &plan_));
    }
    Basis(const Basis &) = delete;
    Basis &operator=(const Basis &) = delete;
    ~Basis() { tg_plan_destroy(plan_); }

    tg_plan *plan() const { return plan_; }
    size_t complex_size() const { return (size_t)Nx * Ny * (Nz / 2 + 1); }
    size_t real_size() const { return (size_t)Nx * Ny * (Nz + 2); }

    void Dealias(double *A)
    {
        double *v[1] = {A};
        check("tg_dealias", tg_dealias(plan_, v, 1));
    }
    void Satisfy_strong_reality_condition_in_Array(double *A)
    {
        double *v[1] = {A};
        check("tg_strong_reality", tg_strong_reality(plan_, v, 1));
    }
    double Get_total_energy(const double *A)
    {
        double e = 0;
        check("tg_energy", tg_energy(plan_, A, &e));
        return e;
    }
    double Get_total_Sn(const double *A, int n)
    {
        double s = 0;
        check("tg_Sn", tg_Sn(plan_, A, n, &s));
        return s;
    }
    std::vector<double> Compute_shell_spectrum(const double *A, int n, int nbins)
    {
        std::vector<double> Sk(nbins);
        check("tg_shell_spectrum", tg_shell_spectrum(plan_, A, n, Sk.data(), nbins));
        return Sk;
    }

    const int Nx, Ny, Nz;

  protected:
    tg_plan *plan_ = nullptr;
};

class FFF_Basis : public Basis {
  public:
    FFF_Basis(int Nx, int Ny, int Nz, std::array<double, 3> kfactor = {1, 1, 1}, int precision = TG_PREC_FP64,
              int dealias = TG_DEALIAS_REF_CORNER, bool strong_reality = true, int device = 0)
        : Basis(TG_BASIS_FFF, Nx, Ny, Nz, kfactor, precision, dealias, strong_reality ? 1 : 0, device)
    {
    }
    // Universal's transform virtuals (universal.h:89-90)
    void Forward_transform(const double *Ar, double *A) { check("tg_forward", tg_forward(plan_, 0, Ar, A)); }
    void Inverse_transform(const double *A, double *Ar) { check("tg_inverse", tg_inverse(plan_, 0, A, Ar)); }
};

// Default kfactor: the reference's RBC box (pi, pi/sqrt 2, pi/sqrt 2)
// (SFF_pencil.cc:67-72); reality 0 none, 1 strong, 2 weak.
class SFF_Basis : public Basis {
  public:
    SFF_Basis(int Nx, int Ny, int Nz,
              std::array<double, 3> kfactor = {3.14159265358979323846, 2.22144146907918312351,
                                               2.22144146907918312351},
              int precision = TG_PREC_FP64, int dealias = TG_DEALIAS_REF_CORNER, int reality = 2, int device = 0)
        : Basis(TG_BASIS_SFF, Nx, Ny, Nz, kfactor, precision, dealias, reality, device)
    {
    }
    // SFF_pencil_tr.cc:43-67; sincos TG_SC_SIN ('S') or TG_SC_COS ('C')
    void Forward_transform(int sincos, const double *Ar, double *A)
    {
        check("tg_forward", tg_forward(plan_, sincos, Ar, A));
    }
    void Inverse_transform(int sincos, const double *A, double *Ar)
    {
        check("tg_inverse", tg_inverse(plan_, sincos, A, Ar));
    }
    void Satisfy_weak_reality_condition_in_Array(double *A)
    {
        double *v[1] = {A};
        check("tg_weak_reality", tg_weak_reality(plan_, v, 1));
    }
};

class FluidVF {
  public:
    FluidVF(Basis &basis, double dissipation_coefficient)
        : basis(basis), dissipation_coefficient(dissipation_coefficient)
    {
        for (int i = 0; i < 3; i++) {
            cvf[i] = DeviceField(basis.complex_size());
            nlin[i] = DeviceField(basis.complex_size());
        }
    }
    std::array<double *, 3> V() const { return {cvf[0].data(), cvf[1].data(), cvf[2].data()}; }
    std::array<double *, 3> N() const { return {nlin[0].data(), nlin[1].data(), nlin[2].data()}; }
    double Compute_total_energy()
    {
        double e = 0;
        for (auto &f : cvf) e += basis.Get_total_energy(f.data());
        return e;
    }
    // FluidVF::Get_dt (FluidVF.cc:546-572): dt_fixed (dt_option 0) or the CFL
    // step of the dealiased field (dt_option 1; dealiases cvf in place)
    double Get_dt(double courant, double dt_fixed, int dt_option = 1)
    {
        if (dt_option == 0) return dt_fixed;
        auto v = V();
        double dt = 0;
        check("tg_get_dt", tg_get_dt(basis.plan(), v.data(), courant, dt_fixed, &dt, nullptr));
        return dt;
    }
    // CVF::Compute_total_k2energy (Cvf.cc:224-227)
    double Compute_total_k2energy()
    {
        double s = 0;
        for (auto &f : cvf) s += basis.Get_total_Sn(f.data(), 2);
        return 2 * s;
    }
    void Init_cond_spectrum(int kind, double k0, double kmin, double kmax, double energy, uint64_t seed)
    {
        auto v = V();
        check("tg_init_field", tg_init_field(basis.plan(), kind, k0, kmin, kmax, energy, seed, v.data()));
    }

    Basis &basis;
    double dissipation_coefficient;
    // hyper_dissipation_coefficient / _exponent (FluidVF.cc:72-84; exponent 0:
    // none), taken by the RK4 step and the adjoint
    double hyper_dissipation_coefficient = 0.0;
    int hyper_dissipation_exponent = 0;
    std::array<DeviceField, 3> cvf, nlin;
};

class FluidSF {
  public:
    FluidSF(Basis &basis, double diffusion_coefficient)
        : basis(basis), diffusion_coefficient(diffusion_coefficient), F(basis.complex_size()),
          nlin(basis.complex_size())
    {
    }
    double Compute_total_energy() { return basis.Get_total_energy(F.data()); }

    Basis &basis;
    double diffusion_coefficient;
    DeviceField F, nlin;
};

struct Nlin_incompress {
    static void Compute_nlin(FluidVF &U)
    {
        auto v = U.V();
        auto n = U.N();
        check("tg_compute_nlin", tg_compute_nlin(U.basis.plan(), v.data(), n.data()));
    }
    // giver form (compute_nlin_ET_VF.cc:71-87): U.nlin from the giver W and U
    static void Compute_nlin_giver(FluidVF &U, FluidVF &W)
    {
        auto v = U.V();
        auto w = W.V();
        auto n = U.N();
        check("tg_compute_nlin_giver", tg_compute_nlin_giver(U.basis.plan(), v.data(), w.data(), n.data()));
    }
    // MHD (compute_nlin_VF_MHD.cc:71-123): U.nlin and W.nlin
    static void Compute_nlin(FluidVF &U, FluidVF &W)
    {
        auto v = U.V();
        auto w = W.V();
        auto nv = U.N();
        auto nw = W.N();
        check("tg_compute_nlin_mhd", tg_compute_nlin_mhd(U.basis.plan(), v.data(), w.data(), nv.data(), nw.data()));
    }
    // scalar / RBC (compute_nlin_SF.cc:75-142): U.nlin and T.nlin
    static void Compute_nlin(FluidVF &U, FluidSF &T)
    {
        auto v = U.V();
        auto n = U.N();
        check("tg_compute_nlin_rbc", tg_compute_nlin_rbc(U.basis.plan(), v.data(), T.F.data(), n.data(),
                                                         T.nlin.data()));
    }
};

class Time_advance_incompress {
  public:
    Time_advance_incompress(FluidVF &U, double dt) : U(U)
    {
        fluid.nu = U.dissipation_coefficient;
        fluid.dt = dt;
        fluid.nu_hyper = U.hyper_dissipation_coefficient;
        fluid.hyper_exponent = U.hyper_dissipation_exponent;
    }
    void Set_energy_supply(double eps, double r_in, double r_out)
    {
        fluid.force_kind = 1;
        fluid.force_eps = eps;
        fluid.force_r_in = r_in;
        fluid.force_r_out = r_out;
    }
    // Band-limited random forcing (compute_force_random.cc:65-130): energy and
    // helicity supply rates on (r_in, r_out], phases redrawn every `tau` time
    // units (seed + n in the n-th interval, as the reference reseeds)
    void Set_random_forcing(double eps, double helicity, double r_in, double r_out, double tau,
                            unsigned long long seed)
    {
        fluid.force_kind = 2;
        fluid.force_eps = eps;
        fluid.force_helicity = helicity;
        fluid.force_r_in = r_in;
        fluid.force_r_out = r_out;
        fluid.force_tau = tau;
        fluid.force_seed = seed;
        seed0_ = seed;
        interval_ = 0;
        random_next_ = tau;
    }
    // Time_advance_step with integration_scheme RK4 (Time_advance.cc:43-65)
    void Time_advance_step(int nsteps = 1)
    {
        auto v = U.V();
        if (fluid.force_kind == 2) {
            // one step a call: a new random interval may begin at any step
            for (int s = 0; s < nsteps; s++) {
                if (time_ >= random_next_) {
                    interval_++;
                    random_next_ += fluid.force_tau;
                    fluid.force_seed = seed0_ + interval_;
                }
                check("tg_rk4_step", tg_rk4_step(U.basis.plan(), v.data(), &fluid, 1, chained_ ? TG_STEP_CONTINUE : 0));
                chained_ = true;
                time_ += fluid.dt;
            }
            return;
        }
        check("tg_rk4_step", tg_rk4_step(U.basis.plan(), v.data(), &fluid, nsteps, chained_ ? TG_STEP_CONTINUE : 0));
        chained_ = true;
        time_ += fluid.dt * nsteps;
    }
    // call after modifying U.cvf outside the time advance
    void Field_modified() { chained_ = false; }

    tg_fluid fluid{};

  private:
    FluidVF &U;
    bool chained_ = false;
    double time_ = 0.0, random_next_ = 0.0;
    unsigned long long seed0_ = 0, interval_ = 0;
};

// RK4 of U (nu) and the magnetic field W (eta) (RK_VF_MHD.cc:126-190)
class Time_advance_incompress_MHD {
  public:
    Time_advance_incompress_MHD(FluidVF &U, FluidVF &W, double dt) : U(U), W(W)
    {
        fluid.nu = U.dissipation_coefficient;
        fluid.dt = dt;
    }
    void Time_advance_step(int nsteps = 1)
    {
        auto v = U.V();
        auto w = W.V();
        check("tg_rk4_step_mhd",
              tg_rk4_step_mhd(U.basis.plan(), v.data(), w.data(), &fluid, W.dissipation_coefficient, nsteps));
    }

    tg_fluid fluid{};

  private:
    FluidVF &U, &W;
};

// Rayleigh-Benard convection, PRLARGE / ULARGE (nu = sqrt(Pr/Ra),
// kappa = 1/sqrt(Pr Ra), Iscalar_main.cc:66-68): RK4 of U and T (RK_SF.cc:97-215)
class Time_advance_incompress_RBC {
  public:
    Time_advance_incompress_RBC(FluidVF &U, FluidSF &T, double Ra, double Pr, double dt) : U(U), T(T)
    {
        rbc.Ra = Ra;
        rbc.Pr = Pr;
        rbc.dt = dt;
    }
    void Compute_force_TO_rhs()
    {
        auto v = U.V();
        auto n = U.N();
        check("tg_compute_rhs_rbc", tg_compute_rhs_rbc(U.basis.plan(), v.data(), T.F.data(), &rbc, n.data(),
                                                       T.nlin.data()));
    }
    void Time_advance_step(int nsteps = 1)
    {
        auto v = U.V();
        check("tg_rk4_step_rbc", tg_rk4_step_rbc(U.basis.plan(), v.data(), T.F.data(), &rbc, nsteps));
    }

    tg_rbc rbc{};

  private:
    FluidVF &U;
    FluidSF &T;
};

struct Correlation {
    // Correlation.cc:144-160 (PRLARGE / ULARGE)
    static double Get_Nusselt_no(FluidVF &U, FluidSF &T, double Ra, double Pr)
    {
        tg_rbc r{Ra, Pr, 1.0};
        double nu = 0;
        check("tg_nusselt", tg_nusselt(U.basis.plan(), U.cvf[0].data(), T.F.data(), &r, &nu));
        return nu;
    }
    // Correlation.cc:225-230
    static double Get_cross_helicity(FluidVF &U, FluidVF &W)
    {
        auto v = U.V();
        auto w = W.V();
        double hc = 0;
        check("tg_cross_helicity", tg_cross_helicity(U.basis.plan(), v.data(), w.data(), &hc));
        return hc;
    }
};

// Checkpointed adjoint of nsteps unforced RK4 steps (no counterpart in the
// reference): Forward runs the steps in place keeping checkpoints (every
// checkpoint_every steps) and the stage tape of the last segment; Backward
// maps lam = dL/dV_n to dL/dV_0 in place.
class Adjoint_incompress {
  public:
    Adjoint_incompress(FluidVF &U, double dt, int checkpoint_every) : U(U), every(checkpoint_every)
    {
        fluid.nu = U.dissipation_coefficient;
        fluid.dt = dt;
        fluid.nu_hyper = U.hyper_dissipation_coefficient;
        fluid.hyper_exponent = U.hyper_dissipation_exponent;
    }
    ~Adjoint_incompress() { (void)tg_adjoint_free(U.basis.plan()); }
    void Forward(int nsteps)
    {
        auto v = U.V();
        check("tg_adjoint_forward", tg_adjoint_forward(U.basis.plan(), v.data(), &fluid, nsteps, every));
    }
    void Backward(std::array<double *, 3> lam) { check("tg_adjoint_backward", tg_adjoint_backward(U.basis.plan(), lam.data())); }

    tg_fluid fluid{};

  private:
    FluidVF &U;
    int every;
};

class EnergyTr {
  public:
    EnergyTr(Basis &basis, std::vector<double> flux_radii, std::vector<double> shell_radii)
        : basis(basis), flux_radii(std::move(flux_radii)), shell_radii(std::move(shell_radii))
    {
    }
    std::vector<double> Compute_flux(FluidVF &U)
    {
        const int n = (int)flux_radii.size() - 1;
        std::vector<double> flux(n + 1);
        auto v = U.V();
        check("tg_flux", tg_flux(basis.plan(), v.data(), flux_radii.data(), n, flux.data()));
        return flux;
    }
    std::vector<double> Compute_shell_tr(FluidVF &U)
    {
        const int n = (int)shell_radii.size() - 1;
        std::vector<double> S((size_t)(n + 1) * (n + 1));
        auto v = U.V();
        check("tg_shell_to_shell", tg_shell_to_shell(basis.plan(), v.data(), shell_radii.data(), n, S.data()));
        return S;
    }

    // both, for flux spheres equal to the shell radii (tg_energy_transfer)
    void Compute_flux_and_shell_tr(FluidVF &U, std::vector<double> &flux, std::vector<double> &S)
    {
        if (flux_radii != shell_radii) throw std::invalid_argument("flux_radii must equal shell_radii");
        const int n = (int)shell_radii.size() - 1;
        flux.assign(n + 1, 0.0);
        S.assign((size_t)(n + 1) * (n + 1), 0.0);
        auto v = U.V();
        check("tg_energy_transfer", tg_energy_transfer(basis.plan(), v.data(), shell_radii.data(), n, flux.data(),
                                                       S.data()));
    }
    // u< -> b> and b< -> b> fluxes (compute_flux.cc:158-211)
    void Compute_flux(FluidVF &U, FluidVF &W, std::vector<double> &u2b, std::vector<double> &b2b)
    {
        const int n = (int)flux_radii.size() - 1;
        u2b.assign(n + 1, 0.0);
        b2b.assign(n + 1, 0.0);
        auto v = U.V();
        auto w = W.V();
        check("tg_flux_mhd", tg_flux_mhd(basis.plan(), v.data(), w.data(), flux_radii.data(), n, u2b.data(),
                                         b2b.data()));
    }

    Basis &basis;
    std::vector<double> flux_radii, shell_radii;
};

}  // namespace tarang
