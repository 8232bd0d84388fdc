// ifluid_main.cpp -- a Tarang-style fluid driver on one GPU, written against
// the C++ host classes of include/tarang_amd.hpp.
//
// Mirrors the iteration loop of Ifluid_main (Ifluid_main.cc:104-131): time
// advance one RK4 step, then the global output (energy, dissipation; the
// glob.d packet of output_energy.cc:69-120), until the final time.  The
// initial field is the seeded spectrum generator (the reference's is seeded
// by time(0)).
//
// usage: tarang_ifluid N steps dt nu [kind=0] [eps=0] [flux_every=0]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tarang_amd.hpp"

int main(int argc, char **argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s N steps dt nu [kind=0] [eps=0] [flux_every=0]\n", argv[0]);
        return 2;
    }
    const int N = std::atoi(argv[1]);
    const int steps = std::atoi(argv[2]);
    const double dt = std::atof(argv[3]);
    const double nu = std::atof(argv[4]);
    const int kind = argc > 5 ? std::atoi(argv[5]) : 0;
    const double eps = argc > 6 ? std::atof(argv[6]) : 0.0;
    const int flux_every = argc > 7 ? std::atoi(argv[7]) : 0;
    try {
        tarang::FFF_Basis basis(N, N, N);
        tarang::FluidVF U(basis, nu);
        U.Init_cond_spectrum(kind, 4.0, 2.0, N / 3.0, 0.5, 1);
        tarang::Time_advance_incompress ta(U, dt);
        if (eps > 0) ta.Set_energy_supply(eps, 2.0, 3.0);
        std::vector<double> radii = {0, 2, 4, 8, N / 4.0, N / 2.0 - 1, 1e300};
        tarang::EnergyTr et(basis, radii, radii);
        double t = 0;
        std::printf("# t E dissipation\n");
        std::printf("%.6f %.15e %.15e\n", t, U.Compute_total_energy(), nu * U.Compute_total_k2energy());
        for (int it = 1; it <= steps; it++) {
            ta.Time_advance_step(1);
            t += dt;
            const double E = U.Compute_total_energy();
            if (E != E) {
                std::fprintf(stderr, "NaN energy at step %d\n", it);  // Ifluid_main.cc:127-128
                return 3;
            }
            std::printf("%.6f %.15e %.15e\n", t, E, nu * U.Compute_total_k2energy());
            if (flux_every > 0 && it % flux_every == 0) {
                auto flux = et.Compute_flux(U);
                std::printf("# flux");
                for (double f : flux) std::printf(" %.6e", f);
                std::printf("\n");
            }
        }
    } catch (const tarang::Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
