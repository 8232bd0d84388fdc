// rbc_main.cpp -- a Tarang-style Rayleigh-Benard driver on one GPU, written
// against the C++ host classes of include/tarang_amd.hpp.
//
// Mirrors the RBC iteration loop (time advance one RK4 step of U and T, then
// the Nusselt number and the energies; Ifluid_main.cc:104-131 with the
// scalar's output, Correlation.cc:144-160).  Initial state: u = 0 and theta
// noise of amplitude 1e-3 in the first modes (this driver's own seeded
// generator; the reference's is seeded by time(0)).
//
// usage: tarang_rbc Nx Ny Nz steps dt Ra Pr [seed=1] [out_prefix]
//   with out_prefix: writes <prefix>_T0.bin (initial theta), <prefix>_V.bin
//   (final u_x, u_y, u_z) and <prefix>_T.bin (final theta), complex FP64
//   (Nx, Ny, Nz/2+1) arrays in native byte order.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tarang_amd.hpp"

namespace {

// SplitMix64 -> uniform in [-1, 1)
struct Rng {
    uint64_t s;
    double next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
};

bool write_fields(const std::string &path, const std::vector<const tarang::DeviceField *> &f)
{
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    bool ok = true;
    for (auto *d : f) {
        auto h = d->download();
        ok = ok && std::fwrite(h.data(), sizeof(double), h.size(), fp) == h.size();
    }
    return std::fclose(fp) == 0 && ok;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s Nx Ny Nz steps dt Ra Pr [seed=1] [out_prefix]\n", argv[0]);
        return 2;
    }
    const int Nx = std::atoi(argv[1]), Ny = std::atoi(argv[2]), Nz = std::atoi(argv[3]);
    const int steps = std::atoi(argv[4]);
    const double dt = std::atof(argv[5]), Ra = std::atof(argv[6]), Pr = std::atof(argv[7]);
    const uint64_t seed = argc > 8 ? std::strtoull(argv[8], nullptr, 10) : 1;
    const std::string prefix = argc > 9 ? argv[9] : "";
    try {
        tarang::SFF_Basis basis(Nx, Ny, Nz);
        tarang::FluidVF U(basis, 0.0);
        tarang::FluidSF T(basis, 0.0);
        // theta noise in lx in [1, m), ky in (-m, m), kz in [0, m)
        const int nzc = Nz / 2 + 1;
        const int m = std::min({32, Nx, Ny / 2, nzc});
        std::vector<double> h(2 * basis.complex_size(), 0.0);
        Rng rng{seed};
        for (int lx = 1; lx < m; lx++)
            for (int ly = 0; ly < Ny; ly++) {
                if (!(ly < m || ly > Ny - m)) continue;
                for (int lz = 0; lz < m; lz++) {
                    const size_t i = ((size_t)lx * Ny + ly) * nzc + lz;
                    h[2 * i] = 1e-3 * rng.next();
                    h[2 * i + 1] = 1e-3 * rng.next();
                }
            }
        T.F.upload(h);
        basis.Satisfy_strong_reality_condition_in_Array(T.F.data());
        if (!prefix.empty() && !write_fields(prefix + "_T0.bin", {&T.F})) {
            std::fprintf(stderr, "cannot write %s_T0.bin\n", prefix.c_str());
            return 4;
        }
        tarang::Time_advance_incompress_RBC ta(U, T, Ra, Pr, dt);
        std::printf("# step t Nu E_u E_theta\n");
        double t = 0;
        for (int it = 1; it <= steps; it++) {
            ta.Time_advance_step(1);
            t += dt;
            const double Eu = U.Compute_total_energy(), Et = T.Compute_total_energy();
            if (Eu != Eu || Et != Et) {
                std::fprintf(stderr, "NaN energy at step %d\n", it);
                return 3;
            }
            std::printf("%d %.6f %.15e %.15e %.15e\n", it, t, tarang::Correlation::Get_Nusselt_no(U, T, Ra, Pr), Eu,
                        Et);
        }
        if (!prefix.empty() && !(write_fields(prefix + "_V.bin", {&U.cvf[0], &U.cvf[1], &U.cvf[2]}) &&
                                 write_fields(prefix + "_T.bin", {&T.F}))) {
            std::fprintf(stderr, "cannot write the final fields under %s\n", prefix.c_str());
            return 4;
        }
    } catch (const tarang::Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
