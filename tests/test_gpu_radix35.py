"""Transform lengths with radix-3 and radix-5 factors (FP64 plans) against the
FP64 oracle, whose mixed-radix FFT is pinned by brute-force DFTs and numpy
fixtures at these lengths (test_oracle.py).

Every compiled length runs through the pass that takes it: x lines of L
points at (L, 8, 16) (x-inverse, x-forward, the fused NS x pass of a forced
RK4 step), y lines at (8, L, 16), z lines of 2M real points at (8, 8, 2M).
Lines of 3*2^a and 5*2^a points straddle wavefronts (their exchanges take
block barriers), and the radix-3/5 passes run through LDS before the
radix-8/4/2 passes.  Cubic grids then carry the other paths: NS with the
2/3 rule, MHD, SFF Rayleigh-Benard, ChFF, the adjoint right-hand side and
the energy-transfer diagnostics."""
import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2, rel_l2_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402
from tarang_amd._lib import TarangGpuError  # noqa: E402

TOL = 1e-12
ODD_L = [24, 40, 48, 80, 96, 120, 160, 192, 240, 320, 384, 480, 640, 768, 960]
ODD_M = [24, 40, 48, 80, 96, 120, 160, 192, 240, 320, 384, 480]


def _transforms(N):
    o = to.Oracle(N, dealias=to.DEALIAS_NONE)
    b = api.FFF_Basis(N, dealias=api.DEALIAS_NONE)
    r = np.random.default_rng(1).standard_normal(o.rshape)
    r[:, :, -2:] = 0.0
    A = b.complex_array()
    b.Forward_transform(dev(r), A)
    e_f = rel_l2(host(A), o.forward(r))
    # the oracle's transform against numpy's (an independent FFT)
    ref_np = np.fft.rfftn(r[:, :, :N[2]], axes=(0, 1, 2)) / (N[0] * N[1] * N[2])
    assert rel_l2(o.forward(r), ref_np) < 1e-14
    rng = np.random.default_rng(2)
    S = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    Sd = dev(S)
    R = b.real_array()
    b.Inverse_transform(Sd, R)
    e_i = rel_l2(host(R), o.inverse(S))
    assert np.array_equal(host(Sd), S)
    return e_f, e_i


def _ns_step(N, dealias=to.DEALIAS_REF_CORNER, strong=True, seed=31):
    o = to.Oracle(N, dealias=dealias, strong_reality=strong)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=max(N) / 3, energy=0.5, seed=seed)
    b = api.FFF_Basis(N, dealias=dealias, strong_reality=strong)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    ref = o.compute_nlin([v.copy() for v in V])
    api.Nlin_incompress.Compute_nlin(U)
    e_n = rel_l2_vec([host(n) for n in U.nlin], ref)
    U.set_field([dev(v) for v in V])
    force = (1, 0.1, 1.0, 3.0)
    ta = api.Time_advance_incompress(U, 1e-3, *force)
    ta.Time_advance_step(2)
    fl = to.Fluid(0.01, 1e-3, *force)
    o.rk4_step(fl, V)
    o.rk4_step(fl, V)
    e_s = rel_l2_vec([host(v) for v in U.cvf], V)
    return e_n, e_s


@pytest.mark.parametrize("L", ODD_L)
def test_x_and_y_lines(L):
    log(f"test_x_and_y_lines L={L}")
    e_f, e_i = _transforms((L, 8, 16))
    e_n, e_s = _ns_step((L, 8, 16))
    e_fy, e_iy = _transforms((8, L, 16))
    log(f"  x: forward {e_f:.2e} inverse {e_i:.2e} nlin {e_n:.2e} 2 RK4 steps {e_s:.2e}; "
        f"y: forward {e_fy:.2e} inverse {e_iy:.2e}")
    assert max(e_f, e_i, e_n, e_s, e_fy, e_iy) < TOL


@pytest.mark.parametrize("M", ODD_M)
def test_z_lines(M):
    log(f"test_z_lines Nz={2 * M}")
    N = (8, 8, 2 * M)
    e_f, e_i = _transforms(N)
    e_n, e_s = _ns_step(N)
    log(f"  forward {e_f:.2e} inverse {e_i:.2e} nlin {e_n:.2e} 2 RK4 steps {e_s:.2e}")
    assert max(e_f, e_i, e_n, e_s) < TOL


@pytest.mark.parametrize("N", [(48, 40, 48), (24, 48, 80)])
def test_ns_twothirds_and_transfer(N):
    log(f"test_ns_twothirds_and_transfer N={N}")
    e_n, e_s = _ns_step(N, dealias=to.DEALIAS_TWOTHIRDS, seed=32)
    o = to.Oracle(N, dealias=to.DEALIAS_TWOTHIRDS, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=6.0, energy=0.5, seed=41)
    for v in V:
        o.dealias(v)
    b = api.FFF_Basis(N, dealias=api.DEALIAS_TWOTHIRDS, strong_reality=True)
    radii = np.array([0.0, 2.0, 4.0, 6.0, 8.0, np.inf])
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    flux, S = api.EnergyTr(b, radii, radii).Compute_flux_and_shell_tr(U)
    S_ref = o.shell_to_shell([v.copy() for v in V], radii)
    flux_ref = o.flux([v.copy() for v in V], radii)
    scale = np.abs(S).sum()
    log(f"  nlin {e_n:.2e} steps {e_s:.2e}; S vs oracle {rel_l2(S, S_ref):.2e}, flux {rel_l2(flux, flux_ref):.2e}, "
        f"sum S / sum|S| {abs(S.sum()) / scale:.1e}")
    assert e_n < TOL and e_s < TOL
    assert rel_l2(S, S_ref) < 1e-11 and rel_l2(flux, flux_ref) < 1e-11
    assert abs(S.sum()) <= 1e-13 * scale and abs(flux[-1]) <= 1e-13 * scale
    Sk = b.Compute_shell_spectrum(U.cvf[0], 0)
    assert rel_l2(Sk, o.shell_spectrum(V[0], 0)) < TOL


def test_mhd_steps():
    N = (24, 40, 48)
    log(f"test_mhd_steps N={N}")
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=8.0, energy=0.5, seed=33)
    W = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=8.0, energy=0.25, seed=34)
    nu, eta, dt = 5e-3, 3e-3, 1e-3
    b = api.FFF_Basis(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    U, B = api.FluidVF(b, nu), api.FluidVF(b, eta)
    U.set_field([dev(v) for v in V])
    B.set_field([dev(w) for w in W])
    ta = api.Time_advance_incompress_MHD(U, B, dt)
    fl = to.Fluid(nu, dt, 0, 0, 0, 0)
    for s in range(2):
        o.rk4_step_mhd(fl, eta, V, W)
        ta.Time_advance_step(1)
        e = rel_l2_vec([host(v) for v in U.cvf] + [host(w) for w in B.cvf], V + W)
        log(f"  MHD step {s + 1}: rel L2 {e:.2e}")
        assert e < TOL


def test_rbc_steps():
    N = (48, 24, 80)
    log(f"test_rbc_steps N={N}")
    kf = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))
    o = to.Oracle(N, kfactor=kf, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    rng = np.random.default_rng(50)
    lx = np.arange(N[0])[:, None, None] * kf[0]
    ly = np.arange(N[1])
    ky = np.where(ly <= N[1] // 2, ly, ly - N[1])[None, :, None] * kf[1]
    kz = np.arange(N[2] // 2 + 1)[None, None, :] * kf[2]
    damp = 0.1 * np.exp(-(np.sqrt(lx ** 2 + ky ** 2 + kz ** 2) / (8.0 * kf[0])) ** 2)
    V = [damp * (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)) for _ in range(3)]
    T = damp * (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape))
    V[0][0] = 0
    T[0] = 0
    Ra, Pr, dt = 1e5, 1.0, 1e-4
    b = api.SFF_Basis(N, kfactor=kf, dealias=api.DEALIAS_REF_CORNER, reality=2)
    U, S = api.FluidVF(b), api.FluidSF(b)
    U.set_field([dev(v) for v in V])
    S.set_field(dev(T))
    ta = api.Time_advance_incompress_RBC(U, S, Ra, Pr, dt)
    ta.Time_advance_step(2)
    for _ in range(2):
        o.rk4_step_rbc(to.Rbc(Ra, Pr, dt, 2), V, T)
    e = rel_l2_vec([host(v) for v in U.cvf] + [host(S.F)], V + [T])
    log(f"  RBC 2 RK4 steps: rel L2 {e:.2e}")
    assert e < TOL


def test_chff_bdf1():
    N = (48, 24, 48)
    kf = (1.0, 0.5, 2.0)
    log(f"test_chff_bdf1 N={N}")
    o = to.Oracle(N, kfactor=kf, dealias=to.DEALIAS_NONE, strong_reality=False)
    rng = np.random.default_rng(60)
    damp = np.exp(-np.arange(N[0]) / 3.0)[:, None, None]
    V = [1e-3 * damp * (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)) for _ in range(3)]
    for v in V:
        v[:, :, 0] = 0.5 * (v[:, :, 0] + np.conj(v[:, (-np.arange(N[1])) % N[1], 0]))
    b = api.ChFF_Basis(N, kfactor=kf)
    U = api.FluidVF(b, 0.05)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress_BDF1(U, 1e-3)
    ta.Time_advance_step(2)
    for _ in range(2):
        P = o.bdf1_chff(0.05, 1e-3, V)
    e = rel_l2_vec([host(v) for v in U.cvf], V)
    ep = rel_l2(host(ta.P), P)
    log(f"  ChFF 2 BDF1 steps: velocity {e:.2e} pressure {ep:.2e}")
    assert e < 1e-11 and ep < 1e-10


def test_adjoint_rhs_vjp():
    N = (24, 40, 48)
    log(f"test_adjoint_rhs_vjp N={N}")
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    V = o.init_field(kind=1, kmin=1.0, kmax=8.0, energy=0.5, seed=31)
    rng = np.random.default_rng(32)
    Lam = [0.1 * (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)) for _ in range(3)]
    ref = o.rhs_vjp(V, Lam)
    b = api.FFF_Basis(N, dealias=api.DEALIAS_REF_CORNER, strong_reality=False)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    out = api.Adjoint_incompress(U, 1e-3).Rhs_vjp([dev(x) for x in Lam])
    e = rel_l2_vec([host(x) for x in out], ref)
    log(f"  rhs_vjp rel L2 {e:.2e}")
    assert e < TOL


def test_fp32_takes_powers_of_two_only():
    with pytest.raises(TarangGpuError):
        api.FFF_Basis((48, 48, 48), precision=api.PREC_FP32)
    with pytest.raises(TarangGpuError):
        api.FFF_Basis((56, 48, 48))      # 7 * 8: no radix-7 kernels


BIG_ODD_L = [192, 384, 768, 960]
BIG_ODD_M = [192, 384, 480]


def _mhd_two_steps(N):
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=max(N) / 3, energy=0.5, seed=35)
    W = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=max(N) / 3, energy=0.25, seed=36)
    b = api.FFF_Basis(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    U, B = api.FluidVF(b, 5e-3), api.FluidVF(b, 3e-3)
    U.set_field([dev(v) for v in V])
    B.set_field([dev(w) for w in W])
    api.Time_advance_incompress_MHD(U, B, 1e-3).Time_advance_step(2)
    fl = to.Fluid(5e-3, 1e-3, 0, 0, 0, 0)
    for _ in range(2):
        o.rk4_step_mhd(fl, 3e-3, V, W)
    return rel_l2_vec([host(v) for v in U.cvf] + [host(w) for w in B.cvf], V + W)


def _rbc_two_steps(N):
    kf = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))
    o = to.Oracle(N, kfactor=kf, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    rng = np.random.default_rng(52)
    lx = np.arange(N[0])[:, None, None] * kf[0]
    ly = np.arange(N[1])
    ky = np.where(ly <= N[1] // 2, ly, ly - N[1])[None, :, None] * kf[1]
    kz = np.arange(N[2] // 2 + 1)[None, None, :] * kf[2]
    damp = 0.1 * np.exp(-(np.sqrt(lx ** 2 + ky ** 2 + kz ** 2) / (40.0 * kf[0])) ** 2)
    V = [damp * (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)) for _ in range(3)]
    T = damp * (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape))
    V[0][0] = 0
    T[0] = 0
    b = api.SFF_Basis(N, kfactor=kf, dealias=api.DEALIAS_REF_CORNER, reality=2)
    U, S = api.FluidVF(b), api.FluidSF(b)
    U.set_field([dev(v) for v in V])
    S.set_field(dev(T))
    ta = api.Time_advance_incompress_RBC(U, S, 1e7, 1.0, 1e-4)
    ta.Compute_nlin()
    Nref, NTref = o.compute_nlin_rbc([v.copy() for v in V], T.copy())
    e_n = rel_l2_vec([host(n) for n in U.nlin] + [host(S.nlin)], Nref + [NTref])
    U.set_field([dev(v) for v in V])
    S.set_field(dev(T))
    api.Time_advance_incompress_RBC(U, S, 1e7, 1.0, 1e-4).Time_advance_step(2)
    for _ in range(2):
        o.rk4_step_rbc(to.Rbc(1e7, 1.0, 1e-4, 2), V, T)
    e = rel_l2_vec([host(v) for v in U.cvf] + [host(S.F)], V + [T])
    log(f"  RBC {N}: Compute_nlin {e_n:.2e}, 2 RK4 steps {e:.2e}")
    return e


def _vjp(N):
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    V = o.init_field(kind=1, kmin=1.0, kmax=max(N) / 3, energy=0.5, seed=37)
    rng = np.random.default_rng(38)
    Lam = [0.1 * (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)) for _ in range(3)]
    ref = o.rhs_vjp(V, Lam)
    b = api.FFF_Basis(N, dealias=api.DEALIAS_REF_CORNER, strong_reality=False)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    out = api.Adjoint_incompress(U, 1e-3).Rhs_vjp([dev(x) for x in Lam])
    return rel_l2_vec([host(x) for x in out], ref)


@pytest.mark.parametrize("L", BIG_ODD_L)
def test_large_odd_x_lines_other_paths(L):
    """The MHD, RBC (sin/cos) and adjoint x passes at the long radix-3/5 x
    lines (their own tile widths and staging paths)."""
    N = (L, 8, 16)
    em, er, ea = _mhd_two_steps(N), _rbc_two_steps(N), _vjp(N)
    # The sin/cos x lines reach K = pi L (3016 at L = 960): the x-derivatives
    # of the nonlinear term multiply the round-off of the high-K modes
    # (which the corner box does not dealias) by K at every stage, and the
    # second step amplifies it again; measured GPU-oracle differences of
    # the state: 1e-15 (L = 384), 1.3e-13 (768), 1.6e-12 (960), largest at
    # lx > 2L/3 where the state is 1e-15 of its peak, while the transforms
    # alone agree with scipy to 5e-16 and Compute_nlin to 3e-14.  The bound
    # scales with K_max^2.
    tol_rbc = TOL * max(1.0, (np.pi * L / 1000.0) ** 2)
    log(f"test_large_odd_x_lines_other_paths L={L}: MHD {em:.2e} RBC {er:.2e} (bound {tol_rbc:.1e}) adjoint {ea:.2e}")
    assert em < TOL and er < tol_rbc and ea < TOL


@pytest.mark.parametrize("M", BIG_ODD_M)
def test_large_odd_z_lines_other_paths(M):
    """The MHD (giver form), RBC and adjoint z passes at long radix-3/5 z lines."""
    N = (8, 8, 2 * M)
    em, er, ea = _mhd_two_steps(N), _rbc_two_steps(N), _vjp(N)
    log(f"test_large_odd_z_lines_other_paths Nz={2 * M}: MHD {em:.2e} RBC {er:.2e} adjoint {ea:.2e}")
    assert em < TOL and er < TOL and ea < TOL


@pytest.mark.parametrize("L", [384, 768, 960])
def test_sff_long_lines_against_scipy(L):
    """sin/cos transforms of long radix-3/5 x lines against scipy.fft's
    DCT-II / DST-II (the convention pinned in test_oracle_sff.py), for the
    GPU and for the oracle."""
    from scipy import fft as sf
    N = (L, 8, 16)
    kf = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))
    o = to.Oracle(N, kfactor=kf, dealias=to.DEALIAS_NONE, strong_reality=False)
    b = api.SFF_Basis(N, kfactor=kf, dealias=api.DEALIAS_NONE, reality=0)
    r = np.zeros(o.rshape)
    r[:, :, :16] = np.random.default_rng(5).standard_normal((L, 8, 16))
    B = np.fft.fft(np.fft.rfft(r[:, :, :16], axis=2), axis=1) / (8 * 16)
    errs = []
    for par, code in ((to.COS, api.SC_COS), (to.SIN, api.SC_SIN)):
        if par == to.COS:
            ref = sf.dct(B, type=2, axis=0) / (2 * L)
        else:
            Y = sf.dst(B, type=2, axis=0) / (2 * L)
            ref = np.zeros_like(Y)
            ref[1:] = Y[:-1]
        A = b.complex_array()
        b.Forward_transform(dev(r), A, code)
        errs += [rel_l2(host(A), ref), rel_l2(o.sff_forward(par, r), ref)]
    log(f"test_sff_long_lines_against_scipy L={L}: GPU cos {errs[0]:.2e} sin {errs[2]:.2e}; "
        f"oracle cos {errs[1]:.2e} sin {errs[3]:.2e}")
    assert max(errs) < 1e-13
