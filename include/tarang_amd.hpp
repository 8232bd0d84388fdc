// tarang_amd.hpp -- C++ host classes over the C ABI (tarang_gpu.h), shaped
// like the reference's basis / field / solver types so that a Tarang driver
// can switch its hot path to the GPU:
//
//   tarang::FFF_Basis                ~ FFF_PENCIL : Universal (universal.h:48-773,
//                                      FFF_pencil.cc:31-277) plus fftk's plan
//   tarang::FluidVF                  ~ FluidVF (FluidVF.h:73-180): cvf V1..V3 and
//                                      nlin1..3 on the device, dissipation_coefficient
//   tarang::Nlin_incompress          ~ Nlin_incompress::Compute_nlin (Nlin.h:64-112)
//   tarang::Time_advance_incompress  ~ Time_advance_incompress (Time_advance.h:69-187)
//   tarang::EnergyTr                 ~ EnergyTr::Compute_flux / Compute_shell_tr
//                                      (EnergyTr.h:239-273)
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

class FFF_Basis {
  public:
    FFF_Basis(int Nx, int Ny, int Nz, std::array<double, 3> kfactor = {1, 1, 1}, int precision = TG_PREC_FP64,
              int dealias = TG_DEALIAS_REF_CORNER, bool strong_reality = true, int device = 0)
        : Nx(Nx), Ny(Ny), Nz(Nz)
    {
        tg_config c{};
        c.basis = TG_BASIS_FFF;
        c.nx = Nx;
        c.ny = Ny;
        c.nz = Nz;
        for (int i = 0; i < 3; i++) c.kfactor[i] = kfactor[i];
        c.precision = precision;
        c.dealias = dealias;
        c.strong_reality = strong_reality ? 1 : 0;
        c.device = device;
        check("tg_plan_create", tg_plan_create(&c, &plan_));
    }
    FFF_Basis(const FFF_Basis &) = delete;
    FFF_Basis &operator=(const FFF_Basis &) = delete;
    ~FFF_Basis() { tg_plan_destroy(plan_); }

    tg_plan *plan() const { return plan_; }
    size_t complex_size() const { return (size_t)Nx * Ny * (Nz / 2 + 1); }
    size_t real_size() const { return (size_t)Nx * Ny * (Nz + 2); }

    // Universal's transform virtuals (universal.h:89-90)
    void Forward_transform(const double *Ar, double *A) { check("tg_forward", tg_forward(plan_, 0, Ar, A)); }
    void Inverse_transform(const double *A, double *Ar) { check("tg_inverse", tg_inverse(plan_, 0, A, Ar)); }
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

  private:
    tg_plan *plan_ = nullptr;
};

class FluidVF {
  public:
    FluidVF(FFF_Basis &basis, double dissipation_coefficient)
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

    FFF_Basis &basis;
    double dissipation_coefficient;
    std::array<DeviceField, 3> cvf, nlin;
};

struct Nlin_incompress {
    static void Compute_nlin(FluidVF &U)
    {
        auto v = U.V();
        auto n = U.N();
        check("tg_compute_nlin", tg_compute_nlin(U.basis.plan(), v.data(), n.data()));
    }
};

class Time_advance_incompress {
  public:
    Time_advance_incompress(FluidVF &U, double dt) : U(U)
    {
        fluid.nu = U.dissipation_coefficient;
        fluid.dt = dt;
    }
    void Set_energy_supply(double eps, double r_in, double r_out)
    {
        fluid.force_kind = 1;
        fluid.force_eps = eps;
        fluid.force_r_in = r_in;
        fluid.force_r_out = r_out;
    }
    // Time_advance_step with integration_scheme RK4 (Time_advance.cc:43-65)
    void Time_advance_step(int nsteps = 1)
    {
        auto v = U.V();
        check("tg_rk4_step", tg_rk4_step(U.basis.plan(), v.data(), &fluid, nsteps, chained_ ? TG_STEP_CONTINUE : 0));
        chained_ = true;
    }
    // call after modifying U.cvf outside the time advance
    void Field_modified() { chained_ = false; }

    tg_fluid fluid{};

  private:
    FluidVF &U;
    bool chained_ = false;
};

class EnergyTr {
  public:
    EnergyTr(FFF_Basis &basis, std::vector<double> flux_radii, std::vector<double> shell_radii)
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

    FFF_Basis &basis;
    std::vector<double> flux_radii, shell_radii;
};

}  // namespace tarang
