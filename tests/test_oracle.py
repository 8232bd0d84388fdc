"""Pins of the FP64 CPU oracle (oracle/tg_oracle.c) by checks that do not
restate its formulas: numpy golden fixtures, brute-force sums, closed forms
and conservation laws.  CPU only.
"""
import os

import numpy as np
import pytest

import tg_oracle as to

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")


def kgrid(o):
    kx = np.fft.fftfreq(o.nx, 1.0 / o.nx)
    ky = np.fft.fftfreq(o.ny, 1.0 / o.ny)
    # Get_kx: lx <= Nx/2 -> +lx (the Nyquist index is +N/2)
    kx[o.nx // 2] = o.nx // 2
    ky[o.ny // 2] = o.ny // 2
    kz = np.arange(o.nz // 2 + 1)
    return np.meshgrid(kx, ky, kz, indexing="ij")


def real_of(o, f):
    r = np.zeros(o.rshape)
    r[:, :, :o.nz] = f
    return r


def xyz(o):
    x = 2 * np.pi * np.arange(o.nx) / o.nx
    y = 2 * np.pi * np.arange(o.ny) / o.ny
    z = 2 * np.pi * np.arange(o.nz) / o.nz
    return np.meshgrid(x, y, z, indexing="ij")


# ---------------------------------------------------------------- transforms

def test_fft1d_against_brute_force():
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 5, 6, 8, 12, 16, 30, 64, 100):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        j = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(j, j) / n)
        assert np.allclose(to.fft1d(x, -1), W @ x, rtol=0, atol=1e-12 * max(1, np.abs(x).sum()))
        assert np.allclose(to.fft1d(x, +1), W.conj() @ x, rtol=0, atol=1e-12 * max(1, np.abs(x).sum()))


def test_transforms_against_numpy_golden():
    g = np.load(GOLDEN)
    n = 0
    while f"shape_{n}" in g:
        shp = tuple(int(v) for v in g[f"shape_{n}"])
        o = to.Oracle(shp, dealias=to.DEALIAS_NONE)
        A = o.forward(real_of(o, g[f"real_{n}"]))
        assert np.abs(A - g[f"spec_{n}"]).max() <= 1e-15 * np.abs(g[f"spec_{n}"]).max() * 10
        r = o.inverse(g[f"S_{n}"])
        ref = g[f"inv_{n}"]
        assert np.abs(r[:, :, :shp[2]] - ref).max() <= 1e-13 * np.abs(ref).max()
        assert np.all(r[:, :, shp[2]:] == 0)
        n += 1
    assert n >= 4


def test_transform_brute_force_small():
    o = to.Oracle((4, 4, 8), dealias=to.DEALIAS_NONE)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((4, 4, 8))
    A = o.forward(real_of(o, f))
    X, Y, Z = np.meshgrid(np.arange(4), np.arange(4), np.arange(8), indexing="ij")
    for lx in range(4):
        for ly in range(4):
            for lz in range(5):
                ph = np.exp(-2j * np.pi * (lx * X / 4 + ly * Y / 4 + lz * Z / 8))
                assert abs(A[lx, ly, lz] - (f * ph).sum() / 128) < 1e-14


def test_single_mode_and_parseval():
    o = to.Oracle((16, 8, 16), dealias=to.DEALIAS_NONE)
    X, Y, Z = xyz(o)
    f = np.cos(3 * X + 2 * Y + 5 * Z) + 0.5 * np.sin(X - 3 * Z)
    A = o.forward(real_of(o, f))
    # cos(k.x) -> 1/2 at (3,2,5); 0.5 sin(x-3z) = 0.5 (e^{i(x-3z)} - e^{-i(x-3z)})/(2i) is
    # stored at k = (-1,0,3) (kz >= 0 half) with coefficient -0.5/(2i)
    assert abs(A[3, 2, 5] - 0.5) < 1e-15
    assert abs(A[-1, 0, 3] - (-0.5 / 2j)) < 1e-15
    nz = np.abs(A) > 1e-12
    assert nz.sum() == 2
    # Parseval on the half spectrum: mean(f^2) = sum_full |A|^2
    w = np.where(np.arange(o.nz // 2 + 1) == 0, 1.0, 2.0)
    w[-1] = 1.0
    assert abs((w * np.abs(A) ** 2).sum() - (f ** 2).mean()) < 1e-14
    r = o.inverse(A)
    assert np.abs(r[:, :, :16] - f).max() < 1e-13


def test_round_trip():
    o = to.Oracle((32, 16, 8), dealias=to.DEALIAS_NONE)
    f = np.random.default_rng(2).standard_normal((32, 16, 8))
    assert np.abs(o.inverse(o.forward(real_of(o, f)))[:, :, :8] - f).max() < 1e-13


# ---------------------------------------------------------------- basis helpers

@pytest.mark.parametrize("N", [16, 24, 32, 64])
def test_dealias_mode_counts(N):
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER)
    A = np.ones(o.cshape, np.complex128)
    o.dealias(A)
    nx = len(range(N // 3 + 1, 2 * N // 3))           # lx in [N/3+1, 2N/3-1]
    nzr = len(range(N // 3 + 1, N // 2 + 1))          # lz in [N/3+1, N/2]
    assert (A == 0).sum() == nx * nx * nzr
    o2 = to.Oracle(N, dealias=to.DEALIAS_TWOTHIRDS)
    A = np.ones(o2.cshape, np.complex128)
    o2.dealias(A)
    m = N // 3
    kept = len([k for k in range(-N // 2 + 1, N // 2 + 1) if abs(k) <= m])
    assert (A != 0).sum() == kept * kept * (m + 1)


def test_strong_reality_makes_inverse_consistent():
    o = to.Oracle(16, dealias=to.DEALIAS_NONE)
    rng = np.random.default_rng(3)
    A = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    o.strong_reality(A)
    assert np.all(A[:, :, -1] == 0)
    # kz=0 plane Hermitian away from the Nyquist row/column
    P = A[:, :, 0]
    for lx in range(16):
        for ly in range(16):
            kx = lx if lx <= 8 else lx - 16
            ky = ly if ly <= 8 else ly - 16
            if abs(kx) == 8 or abs(ky) == 8 or (kx == 0 and ky == 0):
                continue   # Nyquist lines and the mean are not touched by the fill
            assert abs(P[lx, ly] - np.conj(P[-kx % 16, -ky % 16])) < 1e-15
    # forward(inverse(A)) reproduces A away from the Nyquist lines, except
    # that the imaginary part of the mean (k = 0) is dropped by the c2r
    B = o.forward(o.inverse(A))
    KX, KY, KZ = kgrid(o)
    mask = (np.abs(KX) < 8) & (np.abs(KY) < 8) & (KZ < 8)
    mask[0, 0, 0] = False
    assert np.abs((B - A)[mask]).max() < 1e-13
    assert abs(B[0, 0, 0] - A[0, 0, 0].real) < 1e-15


def test_multiplicity_and_radii():
    o = to.Oracle(16)
    assert o.multiplicity(1, 2, 0) == 0.5 and o.multiplicity(1, 2, 3) == 1.0
    assert o.multiplicity(8, 1, 3) == 2.0 and o.multiplicity(1, 8, 0) == 1.0
    assert o.min_radius_outside() == int(np.ceil(np.sqrt(3 * 64)))
    assert o.max_radius_inside() == 8
    r = to.Oracle(512).default_radii(20)
    assert r[0] == 0 and r[1] == 2 and r[3] == 8 and r[18] == 128 and r[19] == 256 and np.isinf(r[20])
    assert abs(r[4] - 8 * 2 ** (4 / 15)) < 1e-12


# ---------------------------------------------------------------- nonlinear term

def _solenoidal(k, rng):
    a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    k = np.asarray(k, float)
    return a - k * (k @ a) / (k @ k)


def _set_mode(A3, k, amp, o):
    kx, ky, kz = k
    if kz < 0 or (kz == 0 and (ky < 0 or (ky == 0 and kx < 0))):
        kx, ky, kz = -kx, -ky, -kz
        amp = np.conj(amp)
    for c in range(3):
        A3[c][kx % o.nx, ky % o.ny, kz] += amp[c]
        if kz == 0:
            A3[c][-kx % o.nx, -ky % o.ny, 0] += np.conj(amp[c])


def test_triad_closed_form():
    """N(k) = i[(k.u_q) u_p + (k.u_p) u_q] for k = p + q (idea of the
    reference's hand check processors/python-scripts/triad.py:21-25)."""
    o = to.Oracle(16, dealias=to.DEALIAS_NONE)
    rng = np.random.default_rng(4)
    p, q = np.array([1, 2, 1]), np.array([2, -1, 1])
    up, uq = _solenoidal(p, rng), _solenoidal(q, rng)
    V = [o.czeros() for _ in range(3)]
    _set_mode(V, p, up, o)
    _set_mode(V, q, uq, o)
    N = o.compute_nlin(V)
    k = p + q
    expect = 1j * ((k @ uq) * up + (k @ up) * uq)
    got = np.array([N[c][k[0] % 16, k[1] % 16, k[2]] for c in range(3)])
    assert np.abs(got - expect).max() < 1e-14
    # p - q interaction (k = p - q: u_q enters conjugated)
    k2 = p - q
    uqc = np.conj(uq)
    expect2 = 1j * ((k2 @ uqc) * up + (k2 @ up) * uqc)
    sgn = k2[2] > 0 or (k2[2] == 0 and (k2[1] > 0 or (k2[1] == 0 and k2[0] > 0)))
    kk = k2 if sgn else -k2
    got2 = np.array([N[c][kk[0] % 16, kk[1] % 16, kk[2]] for c in range(3)])
    assert np.abs(got2 - (expect2 if sgn else np.conj(expect2))).max() < 1e-14


def test_taylor_green_advection():
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    X, Y, Z = xyz(o)
    u = [np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X]
    V = [o.forward(real_of(o, f)) for f in u]
    N = o.compute_nlin(V)
    # (u.grad)u = (sin2x cos^2 z / 2, sin2y cos^2 z / 2, 0)
    ref = [0.5 * np.sin(2 * X) * np.cos(Z) ** 2, 0.5 * np.sin(2 * Y) * np.cos(Z) ** 2, 0 * X]
    for c in range(3):
        assert np.abs(N[c] - o.forward(real_of(o, ref[c]))).max() < 1e-15


def test_beltrami_projected_rhs_vanishes():
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    X, Y, Z = xyz(o)
    A, B, C = 1.0, 0.7, 0.3
    u = [A * np.sin(Z) + C * np.cos(Y), B * np.sin(X) + A * np.cos(Z), C * np.sin(Y) + B * np.cos(X)]
    V = [o.forward(real_of(o, f)) for f in u]
    N = o.compute_nlin(V)
    assert max(np.abs(n).max() for n in N) > 0.1   # (u.grad)u = grad(u^2/2) is not zero
    o.compute_rhs(N, None)
    assert max(np.abs(n).max() for n in N) < 1e-15


def test_rhs_is_solenoidal_and_energy_conserving():
    o = to.Oracle(32, dealias=to.DEALIAS_TWOTHIRDS)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=5)
    o.dealias(V[0]); o.dealias(V[1]); o.dealias(V[2])
    N = o.compute_nlin([v.copy() for v in V])
    o.compute_rhs(N, None)
    KX, KY, KZ = kgrid(o)
    div = KX * N[0] + KY * N[1] + KZ * N[2]
    assert np.abs(div).max() < 1e-14 * max(np.abs(n).max() for n in N)
    # sum_k 2m Re(RHS . V*) = 0 (Galerkin truncation with the 2/3 rule)
    tr = sum(o.shell_mult_single(N[c], V[c], -1.0, np.inf) for c in range(3))
    scale = sum(o.shell_mult_single(np.abs(N[c]).astype(np.complex128), np.abs(V[c]).astype(np.complex128), -1.0, np.inf) for c in range(3))
    assert abs(tr) < 1e-14 * scale


def test_giver_linearity_and_self():
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    V = o.init_field(kind=1, kmin=1.0, kmax=6.0, energy=0.5, seed=6)
    W1 = [o.fill_array_shell(v, 0.0, 3.0) for v in V]
    W2 = [o.fill_array_shell(v, 3.0, np.inf) for v in V]
    Wsum = [a + b for a, b in zip(W1, W2)]
    N1 = o.compute_nlin_giver([v.copy() for v in V], W1)
    N2 = o.compute_nlin_giver([v.copy() for v in V], W2)
    Ns = o.compute_nlin_giver([v.copy() for v in V], Wsum)
    Nself = o.compute_nlin([v.copy() for v in V])
    for c in range(3):
        assert np.abs(N1[c] + N2[c] - Ns[c]).max() < 1e-15
        # the K = 0 mode is excluded from Fill_array_shell (K > 0) and carries no velocity here
        assert np.abs(Ns[c] - Nself[c]).max() < 1e-15


# ---------------------------------------------------------------- time advance

def test_rk4_single_shear_mode_decays_exactly():
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    X, Y, Z = xyz(o)
    u = [0 * X, 0.3 * np.cos(2 * X + Z), 0 * X]       # e_y perpendicular to k = (2,0,1)
    V = [o.forward(real_of(o, f)) for f in u]
    V0 = [v.copy() for v in V]
    fl = to.Fluid(0.05, 1e-2, 0, 0, 0, 0)
    for _ in range(20):
        o.rk4_step(fl, V)
    decay = np.exp(-0.05 * 5 * 0.2)
    assert np.abs(V[1] - decay * V0[1]).max() < 1e-15
    assert np.abs(V[0]).max() < 1e-16 and np.abs(V[2]).max() < 1e-16


@pytest.mark.parametrize("q", [2, 3, 4, 6, 8])
def test_rk4_single_shear_mode_hyperviscous_decay(q):
    """With hyperviscosity (FluidVF hyper_dissipation, Array_exp_ksqr with
    hyper_factor) a single shear mode decays as exp(-(nu K^2 + nu_h K^2
    (K^2)^p) t), p = (q-2)/2 with integer division (q = 4: p = 1); a
    second mode at another |K| checks the K dependence."""
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    X, Y, Z = xyz(o)
    u = [0 * X, 0.3 * np.cos(2 * X + Z) + 0.1 * np.cos(3 * Z), 0 * X]
    V = [o.forward(real_of(o, f)) for f in u]
    V0 = [v.copy() for v in V]
    nu, nuh, dt, n = 0.05, 2e-3, 1e-2, 20
    fl = to.Fluid(nu, dt, 0, 0, 0, 0, nuh, q)
    for _ in range(n):
        o.rk4_step(fl, V)
    p = 1 if q == 4 else (q - 2) // 2
    t = n * dt
    for (lx, ly, lz), K2 in (((2, 0, 1), 5.0), ((0, 0, 3), 9.0)):
        decay = np.exp(-(nu + nuh * K2 ** p) * K2 * t)
        assert abs(V[1][lx, ly, lz] - decay * V0[1][lx, ly, lz]) <= 1e-15 * abs(V0[1][lx, ly, lz]) + 1e-17
    # nu_h = 0 is the plain scheme, bit for bit
    W, W0 = [v.copy() for v in V0], [v.copy() for v in V0]
    o.rk4_step(to.Fluid(nu, dt, 0, 0, 0, 0, 0.0, q), W)
    o.rk4_step(to.Fluid(nu, dt, 0, 0, 0, 0), W0)
    assert all(np.array_equal(a, b) for a, b in zip(W, W0))


@pytest.mark.parametrize("kf", [(1.0, 1.0, 1.0), (1.0, 0.5, 2.0)])
def test_get_dt_cfl_closed_form(kf):
    """FluidVF::Get_dt, dt_option 1: the real-space maxima of single modes
    are their amplitudes (attained at grid points), Delta_x_i = 2 pi /
    (kfactor_i N_i), dt = min(Courant / sum umax_i / Delta_x_i, dt_fixed)."""
    o = to.Oracle((16, 8, 32), kfactor=kf, dealias=to.DEALIAS_REF_CORNER)
    x = 2 * np.pi * np.arange(16) / 16
    y = 2 * np.pi * np.arange(8) / 8
    z = 2 * np.pi * np.arange(32) / 32
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    u = [0.5 * np.cos(X), 0.3 * np.cos(2 * X + Z), -0.2 * np.cos(Y + 3 * Z)]
    V = [o.forward(real_of(o, f)) for f in u]
    dx = [2 * np.pi / kf[i] / n for i, n in enumerate((16, 8, 32))]
    ref = 0.4 / (0.5 / dx[0] + 0.3 / dx[1] + 0.2 / dx[2])
    dt, um = o.get_dt([v.copy() for v in V], 0.4, 1.0)
    assert np.allclose(um, [0.5, 0.3, 0.2], rtol=1e-14, atol=0)
    assert abs(dt - ref) <= 1e-14 * ref
    dt, _ = o.get_dt([v.copy() for v in V], 0.4, 0.5 * ref)      # capped by dt_fixed
    assert dt == 0.5 * ref


def test_get_dt_sff_grid_maxima():
    """SFF: L_x = pi / kfactor_x; the maxima over the cell-centred grid of
    fields built in real space (their transforms are pinned elsewhere)."""
    kf = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))
    N = (16, 8, 16)
    o = to.Oracle(N, kfactor=kf, dealias=to.DEALIAS_NONE, strong_reality=False)
    Lx = np.pi / kf[0]
    x = (np.arange(16) + 0.5) * Lx / 16
    y = 2 * np.pi / kf[1] * np.arange(8) / 8
    z = 2 * np.pi / kf[2] * np.arange(16) / 16
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    u = [0.7 * np.sin(kf[0] * X) * np.cos(kf[2] * Z), 0.4 * np.cos(2 * kf[0] * X), 0.25 * np.cos(kf[1] * Y)]
    par = [to.SIN, to.COS, to.COS]
    V = [o.sff_forward(par[c], real_of(o, u[c])) for c in range(3)]
    dt, um = o.get_dt(V, 0.3, 10.0, sff=True)
    umax = [np.abs(f).max() for f in u]
    assert np.allclose(um, umax, rtol=1e-13, atol=0)
    dx = [Lx / 16, 2 * np.pi / kf[1] / 8, 2 * np.pi / kf[2] / 16]
    ref = 0.3 / (umax[0] / dx[0] + umax[1] / dx[1] + umax[2] / dx[2])
    assert abs(dt - ref) <= 1e-13 * ref


def test_modes_in_shell_brute_force():
    """Universal::Compute_modes_in_shell_array: the modes of the full box
    k_d in (-N_d/2, N_d/2] with ceil(K) = i."""
    o = to.Oracle((16, 8, 16), kfactor=(1.0, 0.5, 2.0))
    c = o.modes_in_shell(6)
    kx = np.arange(-7, 9) * 1.0
    ky = np.arange(-3, 5) * 0.5
    kz = np.arange(-7, 9) * 2.0
    K = np.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    ref = np.bincount(np.ceil(K).astype(int).ravel())[:7]
    assert np.array_equal(c, ref)


@pytest.mark.parametrize("hel", [0.0, 0.6])
def test_random_force_structure(hel):
    """Random band forcing (compute_force_random.cc:65-130): zero outside the
    band, divergence-free, |F|^2 = 2 (F+^2 + F-^2) = 4 e_k / tau, and
    helicity fraction Re(F* . i K x F) / (K |F|^2) = (h_k / K) / e_k
    (the + helical mode is the curl eigenvector of eigenvalue +K); e_k =
    eps / n_shells / modes_in_shell(ceil K); phases set by the seed."""
    N = (16, 16, 16)
    kf = (1.0, 1.0, 1.0)
    o = to.Oracle(N, kfactor=kf)
    eps, r_in, r_out, tau = 0.3, 2.0, 4.0, 0.02
    fl = to.Fluid(0.01, 1e-3, 2, eps, r_in, r_out, 0.0, 0, hel, tau, 17)
    F = o.compute_force(fl, [o.czeros() for _ in range(3)])
    cnt = o.modes_in_shell(4)
    nsh = 4 - 2
    kx = np.fft.fftfreq(16, 1 / 16)[:, None, None] * np.ones((1, 16, 9))
    ky = np.fft.fftfreq(16, 1 / 16)[None, :, None] * np.ones((16, 1, 9))
    kz = np.arange(9)[None, None, :] * np.ones((16, 16, 1))
    K = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    band = (K > r_in) & (K <= r_out)
    F2 = sum(np.abs(f) ** 2 for f in F)
    assert np.all(F2[~band] == 0) and np.all(F2[band] > 0)
    e = eps / nsh / cnt[np.ceil(K[band]).astype(int)]
    h = hel / nsh / cnt[np.ceil(K[band]).astype(int)]
    assert np.allclose(F2[band], 4 * e / tau, rtol=1e-13, atol=0)
    div = kx * F[0] + ky * F[1] + kz * F[2]
    assert np.abs(div).max() <= 1e-13 * np.sqrt(F2.max()) * K.max()
    curl = [1j * (ky * F[2] - kz * F[1]), 1j * (kz * F[0] - kx * F[2]), 1j * (kx * F[1] - ky * F[0])]
    H = sum((np.conj(F[c]) * curl[c]).real for c in range(3))
    frac = H[band] / (K[band] * F2[band])
    assert np.allclose(frac, (h / K[band]) / e, rtol=0, atol=1e-12)
    # the phases: deterministic in the seed, different for another seed
    F_same = o.compute_force(fl, [o.czeros() for _ in range(3)])
    assert all(np.array_equal(a, b) for a, b in zip(F, F_same))
    fl2 = to.Fluid(0.01, 1e-3, 2, eps, r_in, r_out, 0.0, 0, hel, tau, 18)
    F_other = o.compute_force(fl2, [o.czeros() for _ in range(3)])
    assert np.abs(F_other[0] - F[0]).max() > 0.1 * np.abs(F[0]).max()


def test_rk4_fourth_order_convergence():
    """Lawson RK4 is fourth order; with the strong reality condition on, the
    per-step zeroing of the kz = Nz/2 plane (FFF_pencil_basic.cc:175-177)
    makes the scheme first order, which the second half checks."""
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    V0 = o.init_field(kind=0, k0=2.0, energy=0.5, seed=7)

    def run(dt, T=0.2):
        V = [v.copy() for v in V0]
        fl = to.Fluid(0.02, dt, 0, 0, 0, 0)
        for _ in range(int(round(T / dt))):
            o.rk4_step(fl, V)
        return V
    ref = run(0.2 / 128)
    errs = []
    for dt in (0.2 / 4, 0.2 / 8, 0.2 / 16):
        V = run(dt)
        errs.append(np.sqrt(sum(np.sum(np.abs(a - b) ** 2) for a, b in zip(V, ref))))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 13 < r1 < 19 and 13 < r2 < 19, errs
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    ref = run(0.2 / 128)
    e1, e2 = [np.sqrt(sum(np.sum(np.abs(a - b) ** 2) for a, b in zip(run(dt), ref))) for dt in (0.2 / 8, 0.2 / 16)]
    assert 1.7 < e1 / e2 < 2.4


def test_forced_single_mode_energy_closed_form():
    """du/dt = alpha u - nu K^2 u with alpha = eps/(2E): E(t) closed form."""
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    X, Y, Z = xyz(o)
    u = [0 * X, 0.3 * np.cos(2 * X + 2 * Z), 0 * X]   # |K| = 2.83 in (2, 3]
    V = [o.forward(real_of(o, f)) for f in u]
    nu, eps, dt = 0.05, 0.1, 1e-3
    fl = to.Fluid(nu, dt, 1, eps, 2.0, 3.0)
    E0 = sum(o.energy(v) for v in V)
    for _ in range(50):
        o.rk4_step(fl, V)
    K2 = 8.0
    t = 50 * dt
    Einf = eps / (2 * nu * K2)
    E = Einf + (E0 - Einf) * np.exp(-2 * nu * K2 * t)
    assert abs(sum(o.energy(v) for v in V) - E) < 1e-12 * E


# ---------------------------------------------------------------- diagnostics

def test_energy_parseval_and_Sn():
    o = to.Oracle((16, 8, 16), dealias=to.DEALIAS_NONE)
    X, Y, Z = xyz(o)
    f = 0.4 * np.cos(3 * X + 2 * Y + 5 * Z)
    A = o.forward(real_of(o, f))
    assert abs(o.energy(A) - 0.5 * (f ** 2).mean()) < 1e-15
    K2 = 9 + 4 + 25
    # Sn = sum m K^n |A|^2 over the stored half spectrum
    assert abs(o.Sn(A, 2) - K2 * 0.2 ** 2) < 1e-14
    assert abs(o.Sn(A, 0) - 0.2 ** 2) < 1e-15


def test_shell_spectrum_single_mode():
    o = to.Oracle(16)
    X, Y, Z = xyz(o)
    f = 0.4 * np.cos(2 * X + 2 * Y + 1 * Z)            # |K| = 3
    A = o.forward(real_of(o, f))
    Sk = o.shell_spectrum(A, 0)
    assert np.count_nonzero(Sk > 1e-25) == 1 and Sk[3] > 0
    # 2m summed over the modes with ceil(K) = 3, counted by brute force
    cnt = 0.0
    for lx in range(16):
        for ly in range(16):
            for lz in range(9):
                K = o.kmag(lx, ly, lz)
                if int(np.ceil(K)) == 3:
                    cnt += 2 * o.multiplicity(lx, ly, lz)
    assert abs(Sk[3] - 0.2 ** 2 * 9 * 4 * np.pi / cnt) < 1e-15


def test_energy_transfer_identities():
    """2/3 rule, nu = 0: S(n|m) = -S(m|n), sum S = 0, flux from S."""
    o = to.Oracle(32, dealias=to.DEALIAS_TWOTHIRDS)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=8)
    radii = np.array([0.0, 2.0, 4.0, 6.0, 8.0, 10.0, 16.0, np.inf])
    S = o.shell_to_shell([v.copy() for v in V], radii)
    flux = o.flux([v.copy() for v in V], radii)
    ns = len(radii) - 1
    scale = np.abs(S).max()
    assert scale > 0
    Sm = S[1:ns, 1:ns]
    assert np.abs(Sm + Sm.T).max() < 1e-13 * scale
    assert abs(S.sum()) < 1e-13 * scale
    assert abs(flux[-1]) < 1e-14 * scale
    for s in range(1, ns - 1):
        pi = sum(S[m, n] for m in range(1, s + 1) for n in range(s + 1, ns + 1))
        assert abs(flux[s] - pi) < 1e-13 * scale, (s, flux[s], pi)


def test_init_field_properties():
    o = to.Oracle(32)
    for kind in (0, 1):
        V = o.init_field(kind=kind, k0=4.0, kmin=2.0, kmax=10.0, energy=0.5, seed=9)
        assert abs(sum(o.energy(v) for v in V) - 0.5) < 1e-14
        KX, KY, KZ = kgrid(o)
        div = KX * V[0] + KY * V[1] + KZ * V[2]
        assert np.abs(div).max() < 1e-15
        for v in V:
            w = v.copy()
            o.strong_reality(w)
            assert np.abs(w - v).max() < 1e-17


# ---------------------------------------------------------------- MHD

def test_mhd_elsasser_matches_direct_form():
    """N_u = (u.grad)u - (b.grad)b and N_b = (u.grad)b - (b.grad)u, built from
    the giver form, agree with the Elsasser evaluation."""
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    V = o.init_field(kind=1, kmin=1.0, kmax=6.0, energy=0.5, seed=11)
    W = o.init_field(kind=1, kmin=1.0, kmax=6.0, energy=0.25, seed=12)
    NU, NW = o.compute_nlin_mhd([v.copy() for v in V], [w.copy() for w in W])
    uu = o.compute_nlin_giver([v.copy() for v in V], [v.copy() for v in V])
    bb = o.compute_nlin_giver([w.copy() for w in W], [w.copy() for w in W])
    ub = o.compute_nlin_giver([v.copy() for v in V], [w.copy() for w in W])
    bu = o.compute_nlin_giver([w.copy() for w in W], [v.copy() for v in V])
    for c in range(3):
        scale = np.abs(uu[c]).max()
        assert np.abs(NU[c] - (uu[c] - bb[c])).max() < 1e-14 * scale * 10
        assert np.abs(NW[c] - (ub[c] - bu[c])).max() < 1e-14 * scale * 10


def test_mhd_invariants_conserved_by_nonlinear_term():
    o = to.Oracle(32, dealias=to.DEALIAS_TWOTHIRDS)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=13)
    W = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.25, seed=14)
    for c in range(3):
        o.dealias(V[c]); o.dealias(W[c])
    NU, NW = o.compute_nlin_mhd([v.copy() for v in V], [w.copy() for w in W])
    o.compute_rhs(NU, None)
    NW = [-n for n in NW]
    dE = sum(o.shell_mult_single(NU[c], V[c], -1.0, np.inf) + o.shell_mult_single(NW[c], W[c], -1.0, np.inf)
             for c in range(3))
    dHc = sum(o.shell_mult_single(NU[c], W[c], -1.0, np.inf) + o.shell_mult_single(NW[c], V[c], -1.0, np.inf)
              for c in range(3))
    scale = sum(o.shell_mult_single(np.abs(NU[c]).astype(np.complex128), np.abs(V[c]).astype(np.complex128), -1.0, np.inf) for c in range(3))
    assert abs(dE) < 1e-13 * scale and abs(dHc) < 1e-13 * scale
