"""Oracle parity of the kernel instances the 512^3 and 512x1024x1024 step
timings run.

The x and z passes are templated on the line length of their axis, so a thin
grid that keeps that axis at its full size instantiates the same kernel (same
tile width, same launch bounds, same LDS staging and prefetch path) as the
full-size workload, at a size the FP64 oracle finishes in a second:

* NS x pass: k_x_nlin<T, 512, CB> at (512, 8, 16);
* NS z pass: k_z_products<T, 256> (512-point real z lines) at (16, 8, 512);
* RBC x pass: k_x_rbc<T, 512> at SFF (512, 8, 16);
* RBC z pass: k_z_rbc<T, 512> (1024-point real z lines) at SFF (16, 8, 1024).

Each case runs Compute_nlin and one forced (NS) or buoyancy-driven (RBC)
RK4 step in both precisions.  Bars: 1e-12 relative L2 (FP64 transforms),
2e-5 (FP32 transforms).
"""
import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2, rel_l2_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402

TOL = {api.PREC_FP64: 1e-12, api.PREC_FP32: 2e-5}
RBC_KF = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))


@pytest.mark.parametrize("N", [(512, 8, 16), (16, 8, 512), (1024, 8, 16)])
@pytest.mark.parametrize("dealias", [to.DEALIAS_REF_CORNER, to.DEALIAS_TWOTHIRDS])
@pytest.mark.parametrize("prec", [api.PREC_FP64, api.PREC_FP32])
def test_ns_timed_instances(N, dealias, prec):
    """(1024, 8, 16): the x pass at 1024-point lines (twiddles from global
    memory, two-column tiles)."""
    log(f"test_ns_timed_instances N={N} dealias={dealias} prec={prec}")
    # FP32: the rounding of the transforms is scaled by wavenumbers up to N/2
    # in the derivatives (the field holds modes up to N/3)
    tol = TOL[prec] * (max(1, max(N) // 512) if prec == api.PREC_FP32 else 1)
    o = to.Oracle(N, dealias=dealias, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=max(N) / 3, energy=0.5, seed=31)
    b = api.FFF_Basis(N, precision=prec, dealias=dealias, strong_reality=True)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    ref = o.compute_nlin([v.copy() for v in V])
    api.Nlin_incompress.Compute_nlin(U)
    e_n = rel_l2_vec([host(n) for n in U.nlin], ref)
    log(f"  Compute_nlin rel L2 {e_n:.3e}")
    assert e_n < tol
    # one forced step from the same (now dealiased) state
    U.set_field([dev(v) for v in V])
    force = (1, 0.1, 1.0, 3.0)
    fl = to.Fluid(0.01, 1e-3, *force)
    ta = api.Time_advance_incompress(U, 1e-3, *force)
    ta.Time_advance_step(1)
    o.rk4_step(fl, V)
    e_s = rel_l2_vec([host(v) for v in U.cvf], V)
    log(f"  RK4 step rel L2 {e_s:.3e}")
    assert e_s < tol


def _smooth_spec(shape, seed, sine):
    rng = np.random.default_rng(seed)
    nx, ny, nzc = shape
    lx = np.arange(nx)[:, None, None] * RBC_KF[0]
    ly = np.arange(ny)
    ky = np.where(ly <= ny // 2, ly, ly - ny)[None, :, None] * RBC_KF[1]
    kz = np.arange(nzc)[None, None, :] * RBC_KF[2]
    K = np.sqrt(lx ** 2 + ky ** 2 + kz ** 2)
    A = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * np.exp(-(K / (40.0 * RBC_KF[0])) ** 2)
    if sine:
        A[0] = 0
    return (0.1 * A).astype(np.complex128)


@pytest.mark.parametrize("N", [(512, 8, 16), (16, 8, 1024)])
@pytest.mark.parametrize("prec", [api.PREC_FP64, api.PREC_FP32])
def test_rbc_timed_instances(N, prec):
    log(f"test_rbc_timed_instances N={N} prec={prec}")
    tol = TOL[prec]
    Ra, Pr, dt = 1e8, 1.0, 1e-4
    o = to.Oracle(N, kfactor=RBC_KF, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    V = [_smooth_spec(o.cshape, 50 + c, c == 0) for c in range(3)]
    T = _smooth_spec(o.cshape, 53, True)
    b = api.SFF_Basis(N, kfactor=RBC_KF, precision=prec, dealias=api.DEALIAS_REF_CORNER, reality=2)
    U, S = api.FluidVF(b), api.FluidSF(b)
    U.set_field([dev(v) for v in V])
    S.set_field(dev(T))
    ta = api.Time_advance_incompress_RBC(U, S, Ra, Pr, dt)
    Nref, NTref = o.compute_nlin_rbc([v.copy() for v in V], T.copy())
    ta.Compute_nlin()
    e_n = rel_l2_vec([host(n) for n in U.nlin] + [host(S.nlin)], Nref + [NTref])
    log(f"  RBC Compute_nlin rel L2 {e_n:.3e}")
    assert e_n < tol
    U.set_field([dev(v) for v in V])
    S.set_field(dev(T))
    ta = api.Time_advance_incompress_RBC(U, S, Ra, Pr, dt)
    ta.Time_advance_step(1)
    o.rk4_step_rbc(to.Rbc(Ra, Pr, dt, 2), V, T)
    e_s = rel_l2_vec([host(v) for v in U.cvf] + [host(S.F)], V + [T])
    log(f"  RBC RK4 step rel L2 {e_s:.3e}")
    assert e_s < tol


def test_c1_as_written():
    """The CPU-runnable configuration as its text states it: 64^3 FFF NS,
    2/3 dealiasing, RK4 dt = 1e-3, nu = 0.01, 10 steps from the seeded
    K^4 exp(-2 (K/4)^2) field of energy 0.5; energy, dissipation and the
    shell spectrum after every step against the oracle."""
    N, nu, dt = (64, 64, 64), 0.01, 1e-3
    o = to.Oracle(N, dealias=to.DEALIAS_TWOTHIRDS, strong_reality=True)
    V = o.init_field(kind=0, k0=4.0, kmin=1.0, kmax=64.0, energy=0.5, seed=1)
    b = api.FFF_Basis(N, dealias=api.DEALIAS_TWOTHIRDS, strong_reality=True)
    U = api.FluidVF(b, nu)
    U.Init_cond_spectrum(kind=0, k0=4.0, kmin=1.0, kmax=64.0, energy=0.5, seed=1)
    assert rel_l2_vec([host(v) for v in U.cvf], V) < 1e-12
    ta = api.Time_advance_incompress(U, dt)
    fl = to.Fluid(nu, dt, 0, 0.0, 0.0, 0.0)
    for s in range(10):
        o.rk4_step(fl, V)
        ta.Time_advance_step(1)
        E = U.Compute_total_energy()
        eps = U.dissipation()
        Sk = sum(b.Compute_shell_spectrum(U.cvf[c], 0) for c in range(3))
        E_ref = sum(o.energy(v) for v in V)
        eps_ref = nu * 2 * sum(o.Sn(v, 2) for v in V)
        Sk_ref = sum(o.shell_spectrum(v, 0) for v in V)
        e_state = rel_l2_vec([host(v) for v in U.cvf], V)
        log(f"  C1 step {s + 1}: E {E:.15f} eps {eps:.15f} state {e_state:.2e} "
            f"dE {abs(E - E_ref) / E_ref:.1e} deps {abs(eps - eps_ref) / eps_ref:.1e} dSk {rel_l2(Sk, Sk_ref):.1e}")
        assert e_state < 1e-12
        assert abs(E - E_ref) <= 1e-13 * E_ref
        assert abs(eps - eps_ref) <= 1e-13 * eps_ref
        assert rel_l2(Sk, Sk_ref) < 1e-12
    assert E < 0.5  # unforced decay


@pytest.mark.parametrize("N", [(32, 32, 32), (64, 32, 32)])
def test_transfer_zero_sum_twothirds(N):
    """With the 2/3 rule the nonlinear transfer conserves energy exactly in
    every triad: the shell-to-shell matrix is antisymmetric and sums to zero,
    the flux out of a sphere enclosing every excited mode vanishes, and the
    flux equals the transfer from the shells inside to the shells outside,
    all to round-off.  The field is excited for K <= 10, inside the last
    giver shell, so no receiver is cut."""
    o = to.Oracle(N, dealias=to.DEALIAS_TWOTHIRDS, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=10.0, energy=0.5, seed=41)
    for v in V:
        o.dealias(v)
    b = api.FFF_Basis(N, dealias=api.DEALIAS_TWOTHIRDS, strong_reality=True)
    radii = np.array([0.0, 2.0, 4.0, 6.0, 8.0, 10.0, 12.0, 16.0, np.inf])
    ns = len(radii) - 1
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    et = api.EnergyTr(b, radii, radii)
    flux, S = et.Compute_flux_and_shell_tr(U)
    flux_d = et.Compute_flux(U)
    S_d = et.Compute_shell_tr(U)
    S_ref = o.shell_to_shell([v.copy() for v in V], radii)
    scale = np.abs(S).sum()
    Sm = S[1:ns, 1:ns]
    log(f"zero-sum N={N}: sum S {S.sum():.3e} / sum|S| {scale:.3e}; antisym {np.abs(Sm + Sm.T).max():.3e}; "
        f"flux[-1] {flux[-1]:.3e}; max flux {np.abs(flux).max():.3e}; S vs oracle {rel_l2(S, S_ref):.2e}")
    assert scale > 0
    assert abs(S.sum()) <= 1e-13 * scale
    assert np.abs(Sm + Sm.T).max() <= 1e-13 * scale
    assert abs(flux[-1]) <= 1e-13 * scale and abs(flux_d[-1]) <= 1e-13 * scale
    for s in range(1, ns - 1):
        pi = sum(S[m, n] for m in range(1, s + 1) for n in range(s + 1, ns + 1))
        assert abs(flux[s] - pi) <= 1e-13 * scale, (s, flux[s], pi)
    assert np.abs(flux - flux_d).max() <= 1e-13 * scale
    assert rel_l2(S, S_d) < 1e-12 and rel_l2(S, S_ref) < 1e-11


def test_energy_transfer_radii_contract():
    """tg_energy_transfer with a finite last radius that encloses every mode
    equals tg_flux; radii that leave modes outside every giver shell (first
    radius above 0, last radius inside the grid) are rejected."""
    from tarang_amd._lib import ERR_ARG, TarangGpuError
    N = (32, 32, 32)
    o = to.Oracle(N)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=14.0, energy=0.5, seed=43)
    b = api.FFF_Basis(N)
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    rmax = float(b.Min_radius_outside())
    radii = np.array([0.0, 2.0, 4.0, 8.0, 12.0, 16.0, rmax])
    flux, S = api.EnergyTr(b, radii, radii).Compute_flux_and_shell_tr(U)
    flux_d = api.EnergyTr(b, radii, radii).Compute_flux(U)
    flux_ref = o.flux([v.copy() for v in V], radii)
    log(f"finite last radius {rmax}: flux vs tg_flux {np.abs(flux - flux_d).max():.2e}, vs oracle "
        f"{rel_l2(flux, flux_ref):.2e}")
    assert np.abs(flux - flux_d).max() <= 1e-12 * np.abs(flux_d).max()
    assert rel_l2(flux, flux_ref) < 1e-11
    for bad in (np.array([0.5, 2.0, 4.0, np.inf]), np.array([0.0, 2.0, 4.0, 10.0]),
                np.array([0.0, 4.0, 2.0, np.inf])):
        with pytest.raises(TarangGpuError) as ei:
            api.EnergyTr(b, bad, bad).Compute_flux_and_shell_tr(U)
        assert ei.value.code == ERR_ARG


@pytest.mark.parametrize("N", [(512, 8, 16), (16, 8, 512)])
@pytest.mark.parametrize("prec", [api.PREC_FP64, api.PREC_FP32])
def test_mhd_timed_instances(N, prec):
    """The fused MHD x pass at 512-point x lines and the MHD z pass at
    512-point z lines: two RK4 steps against the oracle."""
    log(f"test_mhd_timed_instances N={N} prec={prec}")
    tol = TOL[prec]
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=max(N) / 3, energy=0.5, seed=33)
    W = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=max(N) / 3, energy=0.25, seed=34)
    nu, eta, dt = 5e-3, 3e-3, 1e-3
    b = api.FFF_Basis(N, precision=prec, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    U, B = api.FluidVF(b, nu), api.FluidVF(b, eta)
    U.set_field([dev(v) for v in V])
    B.set_field([dev(w) for w in W])
    ta = api.Time_advance_incompress_MHD(U, B, dt)
    fl = to.Fluid(nu, dt, 0, 0, 0, 0)
    for s in range(2):
        o.rk4_step_mhd(fl, eta, V, W)
        ta.Time_advance_step(1)
        e = rel_l2_vec([host(v) for v in U.cvf] + [host(w) for w in B.cvf], V + W)
        log(f"  MHD step {s + 1}: rel L2 {e:.3e}")
        assert e < tol
