// tg_adjoint.h -- kernels of the adjoint (transpose) of the NS right-hand side
// and of the Lawson RK4 step.
#pragma once
#include "tg_kernels.h"

namespace tg {

// Epilogue of the adjoint x-forward pass (one RK4 stage of the backward sweep;
// r = -c D F(q) is the right-hand side transpose applied to the stage's
// cotangent, lam the cotangent of the step output):
//   ADJ_RHS (-1): lg = r
//   ADJ_S3:       acc = Eh^2 (dt/6 r);   lg = dt/3 Eh lam + dt Eh (dt/6 r)
//   ADJ_S2:       acc += Eh r;           lg = dt/3 Eh lam + dt/2 r
//   ADJ_S1:       acc += Eh r;           lg = dt/6 E1 lam + dt/2 Eh r
//   ADJ_S0:       lam = E1 lam + acc + r
enum { ADJ_RHS = -1, ADJ_S0 = 0, ADJ_S1 = 1, ADJ_S2 = 2, ADJ_S3 = 3 };

struct AdjParams {
    int mode;
    double dt, fh, f1;  // fh = -nu*dt/2, f1 = -nu*dt
    double fhh, f1h;    // the same with nu_hyper (hyperviscosity)
    int hq;             // hyper_exponent q (0: none)
    double scale;       // 1/(Nx Ny Nz)
};

// x-inverse of the symmetrised fields (-i K_m M_j - i K_j M_m)/c with
// M = P G, c = 1 on kz = 0 and kz = Nz/2, 2 elsewhere: the pairs xx, yy, xy,
// yz, zx (the field is trace-free, zz = -xx - yy, as K.M = 0).  G: FP64
// boundary layout; S: five T work fields.
template <typename T>
hipError_t launch_xinv_sym(const Grid &g, double *const *G, void *const *S, const void *twx, hipStream_t st);

// z pass: c2r of the five symmetrised lines, q_m = sum_j u_j sigma_jm against
// the taped velocity lines (sigma_zz = -sigma_xx - sigma_yy), r2c of q (three
// T work fields).
template <typename T>
hipError_t launch_z_adj(const Grid &g, void *const *S, void *const *Q, void *const *tape, const void *twm,
                        const void *twn, hipStream_t st);

// x-forward of q with r = -c D scale F(q) and the stage epilogue above.
// With S_next (modes S3..S0), the next stage's cotangent (lg, or lam after
// S0) is not stored as lg but taken, in registers, through the next stage's
// launch_xinv_sym work into S_next.
template <typename T>
hipError_t launch_xfwd_adj(const Grid &g, const AdjParams &ap, void *const *Q, double *const *lam,
                           double *const *acc, double *const *lg, const void *twx, hipStream_t st,
                           void *const *S_next = nullptr);

}  // namespace tg
