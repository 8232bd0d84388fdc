"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle.

Bar: FP64 transform path <= 1e-12 relative L2 error against the oracle on
identical seeded inputs; FP32 transform path <= 1e-5 (transforms) and the
tolerances stated per test.
"""
import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2, rel_l2_vec

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from tarang_amd import api  # noqa: E402

FP64_TOL = 1e-12


def _rng_real(shape, seed):
    r = np.random.default_rng(seed).standard_normal(shape)
    r[:, :, -2:] = 0.0
    return r


@pytest.mark.parametrize("N", [(16, 16, 16), (32, 16, 64), (64, 64, 64), (8, 32, 16), (8, 512, 16), (8, 1024, 16)])
@pytest.mark.parametrize("prec", [api.PREC_FP64, api.PREC_FP32])
def test_transforms(N, prec):
    """(8, 512, 16) and (8, 1024, 16): y lines of 512 and 1024 points (the
    1024-thread FP32 y tiles of 16 and 8 columns)."""
    log(f"test_transforms N={N} prec={prec}")
    o = to.Oracle(N, dealias=to.DEALIAS_NONE)
    b = api.FFF_Basis(N, precision=prec, dealias=api.DEALIAS_NONE)
    r = _rng_real(o.rshape, 1)
    spec_ref = o.forward(r)
    Ar = dev(r)
    A = b.complex_array()
    b.Forward_transform(Ar, A)
    e_f = rel_l2(host(A), spec_ref)
    # inverse of a non-Hermitian random spectrum (imag parts of kz=0 and Nyquist ignored)
    rng = np.random.default_rng(2)
    S = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    real_ref = o.inverse(S)
    Sd = dev(S)
    Br = b.real_array()
    b.Inverse_transform(Sd, Br)
    e_i = rel_l2(host(Br), real_ref)
    assert np.array_equal(host(Sd), S), "inverse must preserve its input"
    tol = FP64_TOL if prec == api.PREC_FP64 else 1e-5
    log(f"  forward rel L2 {e_f:.3e}  inverse rel L2 {e_i:.3e}")
    assert e_f < tol and e_i < tol


def _init(o, kind=0, seed=1, energy=0.5, kmax=None):
    return o.init_field(kind=kind, k0=4.0, kmin=2.0, kmax=kmax if kmax else o.nx / 3, energy=energy, seed=seed)


@pytest.mark.parametrize("N", [(16, 16, 16), (32, 32, 32), (32, 16, 64)])
@pytest.mark.parametrize("dealias", [to.DEALIAS_REF_CORNER, to.DEALIAS_TWOTHIRDS, to.DEALIAS_NONE])
def test_compute_nlin(N, dealias):
    log(f"test_compute_nlin N={N} dealias={dealias}")
    o = to.Oracle(N, dealias=dealias)
    V = _init(o, kind=1, seed=3)
    b = api.FFF_Basis(N, dealias=dealias)
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    ref = o.compute_nlin(V)
    api.Nlin_incompress.Compute_nlin(U)
    e = rel_l2_vec([host(n) for n in U.nlin], ref)
    ev = rel_l2_vec([host(v) for v in U.cvf], V)  # state dealiased in place
    log(f"  nlin rel L2 {e:.3e}  state {ev:.3e}")
    assert e < FP64_TOL and ev < FP64_TOL


def test_compute_rhs_forced():
    N = (32, 32, 32)
    o = to.Oracle(N)
    V = _init(o, kind=1, seed=4)
    fl = to.Fluid(0.01, 1e-3, 1, 0.1, 2.0, 3.0)
    ref = o.compute_nlin([v.copy() for v in V])
    F = o.compute_force(fl, V)
    o.compute_rhs(ref, F)
    b = api.FFF_Basis(N)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, 1e-3, 1, 0.1, 2.0, 3.0)
    ta.Compute_force_TO_rhs()
    e = rel_l2_vec([host(n) for n in U.nlin], ref)
    log(f"test_compute_rhs_forced rel L2 {e:.3e}")
    assert e < FP64_TOL


@pytest.mark.parametrize("prec,tol", [(api.PREC_FP64, 1e-13), (api.PREC_FP32, 1e-5)])
def test_rhs_trace_shift_is_a_projected_gradient(prec, tol):
    """The projected right-hand side (five trace-shifted products) equals
    -P N with N from Compute_nlin (the reference's six products), both on the
    GPU: the shift changes N by a gradient only, which P removes mode by mode.
    The projection here is numpy's, from the same wavenumbers."""
    N = (64, 32, 64)
    log(f"test_rhs_trace_shift_is_a_projected_gradient prec={prec}")
    o = to.Oracle(N)
    V = _init(o, kind=1, seed=14)
    b = api.FFF_Basis(N, precision=prec)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    api.Nlin_incompress.Compute_nlin(U)
    Nl = [host(n) for n in U.nlin]
    U.set_field([dev(v) for v in V])
    api.Time_advance_incompress(U, 1e-3).Compute_force_TO_rhs()
    R = [host(n) for n in U.nlin]
    # the reference's wavenumbers: +N/2 at the Nyquist index (get_kx/get_ky), which
    # matters in K.N (numpy's fftfreq takes -N/2 there)
    kx = np.array([l if l <= N[0] // 2 else l - N[0] for l in range(N[0])], float)[:, None, None]
    ky = np.array([l if l <= N[1] // 2 else l - N[1] for l in range(N[1])], float)[None, :, None]
    kz = np.arange(N[2] // 2 + 1)[None, None, :]
    K = [kx + 0 * ky + 0 * kz, ky + 0 * kx + 0 * kz, kz + 0 * kx + 0 * ky]
    K2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    div = sum(K[i] * Nl[i] for i in range(3))
    with np.errstate(invalid="ignore", divide="ignore"):
        q = np.where(K2 > 0, div / K2, 0.0)
    ref = [-(Nl[i] - K[i] * q) for i in range(3)]
    e = rel_l2_vec(R, ref)
    shift = rel_l2_vec(R, [-n for n in Nl])
    log(f"  RHS vs -P N: rel L2 {e:.3e} (vs -N unprojected: {shift:.3e})")
    assert e < tol and shift > 1e-3


def test_compute_nlin_giver():
    N = (32, 32, 32)
    o = to.Oracle(N)
    V = _init(o, kind=1, seed=5)
    W = [o.fill_array_shell(v, 2.0, 6.0) for v in V]
    b = api.FFF_Basis(N)
    U = api.FluidVF(b)
    Wf = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    Wf.set_field([dev(w) for w in W])
    ref = o.compute_nlin_giver([v.copy() for v in V], [w.copy() for w in W])
    api.Nlin_incompress.Compute_nlin(U, Wf)
    e = rel_l2_vec([host(n) for n in U.nlin], ref)
    log(f"test_compute_nlin_giver rel L2 {e:.3e}")
    assert e < FP64_TOL


@pytest.mark.parametrize("case", ["c1_64", "forced_32", "twothirds_noreal_32"])
def test_rk4_steps(case):
    if case == "c1_64":
        N, nu, dt, force, steps, dealias, real = (64, 64, 64), 0.01, 1e-3, (0, 0.0, 0.0, 0.0), 3, to.DEALIAS_REF_CORNER, True
        kind = 0
    elif case == "forced_32":
        N, nu, dt, force, steps, dealias, real = (32, 32, 32), 0.005, 2e-3, (1, 0.1, 2.0, 3.0), 3, to.DEALIAS_REF_CORNER, True
        kind = 1
    else:
        N, nu, dt, force, steps, dealias, real = (32, 16, 64), 0.02, 1e-3, (0, 0.0, 0.0, 0.0), 2, to.DEALIAS_TWOTHIRDS, False
        kind = 1
    log(f"test_rk4_steps {case}")
    o = to.Oracle(N, dealias=dealias, strong_reality=real)
    V = _init(o, kind=kind, seed=7)
    fl = to.Fluid(nu, dt, *force)
    b = api.FFF_Basis(N, dealias=dealias, strong_reality=real)
    U = api.FluidVF(b, nu)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, dt, *force)
    for s in range(steps):
        o.rk4_step(fl, V)
        ta.Time_advance_step(1)
        e = rel_l2_vec([host(v) for v in U.cvf], V)
        log(f"  step {s + 1}: rel L2 {e:.3e}")
        assert e < FP64_TOL


@pytest.mark.parametrize("mode", ["one_call", "chained"])
@pytest.mark.parametrize("prec,tol", [(api.PREC_FP64, FP64_TOL), (api.PREC_FP32, 2e-5)])
def test_rk4_steps_reality_planes(mode, prec, tol):
    """Three forced steps, in one call or in three chained calls: the stage-4
    x pass leaves the next step's stage-1 input, and after the strong reality
    condition its kz = 0 and Nz/2 columns (and the forcing band's slots) are
    re-transformed.  The initial kz = 0 planes are perturbed off Hermitian
    symmetry so that the condition changes them."""
    N, nu, dt, force = (32, 16, 32), 0.005, 2e-3, (1, 0.1, 1.0, 3.0)
    log(f"test_rk4_steps_reality_planes mode={mode} prec={prec}")
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=10.0, energy=0.5, seed=81)
    rng = np.random.default_rng(82)
    for c in range(3):
        V[c][:, :, 0] += 1e-3 * (rng.standard_normal(V[c].shape[:2]) + 1j * rng.standard_normal(V[c].shape[:2]))
    b = api.FFF_Basis(N, precision=prec, dealias=api.DEALIAS_REF_CORNER, strong_reality=True)
    U = api.FluidVF(b, nu)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, dt, *force)
    fl = to.Fluid(nu, dt, *force)
    for _ in range(3):
        o.rk4_step(fl, V)
    if mode == "one_call":
        ta.Time_advance_step(3)
    else:
        for _ in range(3):
            ta.Time_advance_step(1)
    e = rel_l2_vec([host(v) for v in U.cvf], V)
    log(f"  after 3 steps: rel L2 {e:.3e}")
    assert e < tol


def test_rk4_fifty_steps():
    """Fifty forced RK4 steps at 64^3 (strong reality, corner dealias) in
    chained calls against the oracle's six-product steps: the trace-shifted
    stage stays within rounding over a long run (state and energy)."""
    N, nu, dt, force = (64, 64, 64), 0.005, 2e-3, (1, 0.1, 2.0, 3.0)
    log("test_rk4_fifty_steps")
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = _init(o, kind=1, seed=17)
    b = api.FFF_Basis(N, dealias=api.DEALIAS_REF_CORNER, strong_reality=True)
    U = api.FluidVF(b, nu)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, dt, *force)
    fl = to.Fluid(nu, dt, *force)
    for s in range(50):
        o.rk4_step(fl, V)
        ta.Time_advance_step(1)
        if (s + 1) % 10 == 0:
            e = rel_l2_vec([host(v) for v in U.cvf], V)
            E, E_ref = U.Compute_total_energy(), sum(o.energy(v) for v in V)
            log(f"  step {s + 1}: rel L2 {e:.3e}, energy {abs(E - E_ref) / E_ref:.3e}")
            assert e < 1e-11 and abs(E - E_ref) < 1e-12 * E_ref


def test_rk4_fp32_vs_fp64():
    N = (32, 32, 32)
    o = to.Oracle(N)
    V = _init(o, kind=1, seed=8)
    fl = to.Fluid(0.01, 1e-3, 0, 0, 0, 0)
    b = api.FFF_Basis(N, precision=api.PREC_FP32)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, 1e-3)
    for s in range(3):
        o.rk4_step(fl, V)
    ta.Time_advance_step(3)
    e = rel_l2_vec([host(v) for v in U.cvf], V)
    E_gpu = U.Compute_total_energy()
    E_ref = sum(o.energy(v) for v in V)
    log(f"test_rk4_fp32 state rel L2 {e:.3e} energy rel {abs(E_gpu - E_ref) / E_ref:.3e}")
    assert e < 1e-5 and abs(E_gpu - E_ref) / E_ref < 1e-6


def test_energy_spectrum():
    N = (32, 32, 32)
    o = to.Oracle(N)
    V = _init(o, kind=0, seed=9)
    b = api.FFF_Basis(N)
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    for c in range(3):
        e_ref = o.energy(V[c])
        e = b.Get_total_energy(U.cvf[c])
        assert abs(e - e_ref) <= 1e-13 * abs(e_ref)
        for n in (0, 2):
            s_ref = o.Sn(V[c], n)
            s = b.Get_total_Sn(U.cvf[c], n)
            assert abs(s - s_ref) <= 1e-13 * abs(s_ref)
        sk_ref = o.shell_spectrum(V[c], 0)
        sk = b.Compute_shell_spectrum(U.cvf[c], 0)
        assert rel_l2(sk, sk_ref) < FP64_TOL


def test_flux_and_shell_to_shell():
    N = (32, 32, 32)
    o = to.Oracle(N)
    V = _init(o, kind=1, seed=10)
    b = api.FFF_Basis(N)
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    radii = np.array([0.0, 2.0, 4.0, 6.0, 8.0, 10.0, 12.0, 16.0, np.inf])
    flux_ref = o.flux([v.copy() for v in V], radii)
    S_ref = o.shell_to_shell([v.copy() for v in V], radii)
    et = api.EnergyTr(b, radii, radii)
    flux = et.Compute_flux(U)
    S = et.Compute_shell_tr(U)
    log(f"flux rel {rel_l2(flux, flux_ref):.3e}  S rel {rel_l2(S, S_ref):.3e}")
    assert rel_l2(flux, flux_ref) < 1e-11
    assert rel_l2(S, S_ref) < 1e-11


@pytest.mark.parametrize("N,radii", [
    ((32, 32, 32), np.array([0.0, 2.0, 4.0, 6.0, 8.0, 10.0, 12.0, 16.0, np.inf])),
    ((64, 32, 32), None),
])
def test_energy_transfer_combined(N, radii):
    """tg_energy_transfer: the shell-to-shell matrix and the flux derived from
    its (uncut) rows against the oracle's direct evaluations.  The last shell
    (largest inscribed radius, inf) holds the corner modes the flux's receiver
    includes."""
    o = to.Oracle(N)
    V = _init(o, kind=1, seed=12)
    b = api.FFF_Basis(N)
    if radii is None:
        radii = b.default_radii(8)
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    flux_ref = o.flux([v.copy() for v in V], radii)
    S_ref = o.shell_to_shell([v.copy() for v in V], radii)
    flux, S = api.EnergyTr(b, radii, radii).Compute_flux_and_shell_tr(U)
    log(f"combined N={N}: flux rel {rel_l2(flux, flux_ref):.3e}  S rel {rel_l2(S, S_ref):.3e}")
    assert rel_l2(flux, flux_ref) < 1e-11
    assert rel_l2(S, S_ref) < 1e-11


def test_init_field():
    N = (32, 32, 32)
    o = to.Oracle(N)
    for kind in (0, 1):
        V = o.init_field(kind=kind, k0=4.0, kmin=2.0, kmax=10.0, energy=0.5, seed=11)
        b = api.FFF_Basis(N)
        U = api.FluidVF(b)
        U.Init_cond_spectrum(kind=kind, k0=4.0, kmin=2.0, kmax=10.0, energy=0.5, seed=11)
        e = rel_l2_vec([host(v) for v in U.cvf], V)
        log(f"test_init_field kind={kind} rel L2 {e:.3e}")
        assert e < FP64_TOL
