"""Pins of the oracle's Chebyshev (ChFF) restatement (oracle/tg_oracle_chff.c)
by numpy's Chebyshev module: transforms of sampled Chebyshev series at the
Chebyshev-Gauss points, the derivative against chebder, the Helmholtz
solver against the tau conditions of u'' - lambda u = -f with its wall
values, and the BDF1 step's wall conditions and continuity.  CPU only."""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest

import tg_oracle as to


def gauss(n):
    return np.cos(np.pi * (np.arange(n) + 0.5) / n)


def test_transform_of_chebyshev_series():
    N = (16, 4, 8)
    o = to.Oracle(N, kfactor=(0.7, 1.0, 1.0), dealias=to.DEALIAS_NONE, strong_reality=False)
    rng = np.random.default_rng(1)
    c0, c1 = rng.standard_normal(16), rng.standard_normal(16)
    x = gauss(16)
    y = 2 * np.pi * np.arange(4) / 4
    r = np.zeros(o.rshape)
    # f = g0(x) + g1(x) cos(y): (ky, kz) = (0, 0) holds c0, (+-1, 0) hold c1 / 2
    r[:, :, :8] = (C.chebval(x, c0)[:, None, None] + C.chebval(x, c1)[:, None, None] * np.cos(y)[None, :, None])
    A = o.chff_forward(r)
    assert np.abs(A[:, 0, 0] - c0).max() < 1e-14
    assert np.abs(A[:, 1, 0] - c1 / 2).max() < 1e-14 and np.abs(A[:, 3, 0] - c1 / 2).max() < 1e-14
    mask = np.ones(A.shape, bool)
    mask[:, [0, 1, 3], 0] = False
    assert np.abs(A[mask]).max() < 1e-14
    back = o.chff_inverse(A)
    assert np.abs(back[:, :, :8] - r[:, :, :8]).max() < 1e-13


def test_xderiv_against_chebder():
    N = (16, 4, 8)
    kf = 0.7
    o = to.Oracle(N, kfactor=(kf, 1.0, 1.0))
    rng = np.random.default_rng(2)
    A = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    B = o.cheb_xderiv(A)
    ref = kf * np.concatenate([C.chebder(A, axis=0), np.zeros((1,) + A.shape[1:])], axis=0)
    assert np.abs(B - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("lam,vp,vm", [(3.7, 0.3, -1.2), (0.0, 0.0, 0.0), (250.0, 1.0, 0.0)])
def test_helmholtz_line_tau_conditions(lam, vp, vm):
    """(D^2 - lambda) u = -f in the tau sense: coefficients 0..N-3 of
    u'' - lambda u + f vanish, u(1) = vp, u(-1) = vm."""
    N = 24
    o = to.Oracle((N, 4, 8))
    f = np.random.default_rng(3).standard_normal(N)
    u = o.helmholtz_line(f, lam, vp, vm)
    res = np.concatenate([C.chebder(u, 2), [0, 0]]) - lam * u + f
    assert np.abs(res[:N - 2]).max() <= 1e-12 * max(1.0, np.abs(f).max())
    assert abs(C.chebval(1.0, u) - vp) < 1e-12 and abs(C.chebval(-1.0, u) - vm) < 1e-12


def test_helmholtz_full_array_lambda_per_line():
    N = (16, 4, 8)
    o = to.Oracle(N, kfactor=(1.0, 0.5, 2.0))
    rng = np.random.default_rng(4)
    R = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    U = o.helmholtz(R, 1.5, 0.25 + 0.5j, -1.0j)
    for ly in range(4):
        for lz in range(5):
            ky = ly if ly <= 2 else ly - 4
            lam = (ky * 0.5) ** 2 + (lz * 2.0) ** 2 + 1.5
            u = U[:, ly, lz]
            res = np.concatenate([C.chebder(u, 2), [0, 0]]) - lam * u + R[:, ly, lz]
            assert np.abs(res[:-2]).max() <= 1e-12 * np.abs(R).max() * max(1, lam)
            assert abs(C.chebval(1.0, u) - (0.25 + 0.5j)) < 1e-12 and abs(C.chebval(-1.0, u) + 1.0j) < 1e-12


def test_bdf1_wall_conditions_and_continuity():
    """From a divergence-free field, after a BDF1 step every component
    vanishes at both walls and the influence-matrix correction makes
    d u_x / dx vanish there too (exact conditions of the scheme).  The
    scheme has no tau correction, so continuity holds up to the truncation
    of its Helmholtz solves, which converges spectrally: bounded here."""
    N = (32, 8, 8)
    kf = (1.0, 1.0, 1.0)
    o = to.Oracle(N, kfactor=kf, dealias=to.DEALIAS_NONE, strong_reality=False)
    rng = np.random.default_rng(5)
    KY = np.fft.fftfreq(8, 1 / 8)
    KY[4] = 4
    # a divergence-free no-slip start, exact in 32 modes: per (ky, kz) line
    # u = a (phi, i Ky phi' / K^2, i Kz phi' / K^2) + b (0, i Kz chi, -i Ky chi)
    # with phi = (1 - x^2)^2 h, chi = (1 - x^2) q, h, q of degree 6
    w = np.array([0.5, 0.0, -0.5])               # 1 - x^2
    V = [np.zeros(o.cshape, complex) for _ in range(3)]
    for ly in range(8):
        for lz in range(5):
            ky, kz = KY[ly], float(lz)
            if ly == 4 or lz == 4 or (ky == 0 and kz == 0):
                continue
            K2 = ky * ky + kz * kz
            h = rng.standard_normal(7) + 1j * rng.standard_normal(7)
            q = rng.standard_normal(7) + 1j * rng.standard_normal(7)
            phi = C.chebmul(C.chebmul(w, w), h)
            chi = C.chebmul(w, q)
            dphi = C.chebder(phi)
            pad = lambda c: np.concatenate([c, np.zeros(32 - len(c))])
            V[0][:, ly, lz] = 1e-2 * pad(phi)
            V[1][:, ly, lz] = 1e-2 * (pad(1j * ky * dphi / K2) + pad(1j * kz * chi))
            V[2][:, ly, lz] = 1e-2 * (pad(1j * kz * dphi / K2) - pad(1j * ky * chi))
    for v in V:
        v[:, :, 0] = 0.5 * (v[:, :, 0] + np.conj(v[:, (-np.arange(8)) % 8, 0]))   # Hermitian kz = 0 plane
    nu, dt = 0.05, 1e-2
    P = o.bdf1_chff(nu, dt, V)
    grad_scale = np.abs(C.chebder(V[0], axis=0)).max()
    worst = 0.0
    for ly in range(8):
        for lz in range(5):
            for c in range(3):
                u = V[c][:, ly, lz]
                assert abs(C.chebval(1.0, u)) < 1e-12 and abs(C.chebval(-1.0, u)) < 1e-12, (ly, lz, c)
            if (ly, lz) != (0, 0):
                du = C.chebder(V[0][:, ly, lz])
                assert abs(C.chebval(1.0, du)) < 1e-10 and abs(C.chebval(-1.0, du)) < 1e-10, (ly, lz)
            div = (np.concatenate([C.chebder(V[0][:, ly, lz]), [0]]) + 1j * KY[ly] * V[1][:, ly, lz]
                   + 1j * lz * V[2][:, ly, lz])
            worst = max(worst, np.abs(div).max())
    # the residual converges spectrally (1e-2 of the gradient at 16 modes, 2e-6 at 32)
    assert worst < 1e-4 * grad_scale, (worst, grad_scale)
    assert np.all(V[0][:, 0, 0] == 0) and np.all(P[:, 0, 0] == 0)


def test_bdf1_mean_flow_line_against_dense_tau_solve():
    """The (ky, kz) = (0, 0) line has no pressure gradient along y, z and
    u_x = 0: each of u_y, u_z solves (nu D^2 - 1/dt) u = N - u_n/dt with
    u(+-1) = 0.  Solved here as a dense tau system with numpy (rows
    0..N-3 of the operator, two boundary rows)."""
    N = (16, 4, 4)
    o = to.Oracle(N, dealias=to.DEALIAS_NONE, strong_reality=False)
    rng = np.random.default_rng(6)
    damp = np.exp(-np.arange(16) / 2.0)
    V = [np.zeros(o.cshape, complex) for _ in range(3)]
    V[1][:, 0, 0] = 1e-3 * damp * rng.standard_normal(16)
    V[2][:, 0, 0] = 1e-3 * damp * rng.standard_normal(16)
    Vn = [v.copy() for v in V]
    NL = o.compute_nlin_chff([v.copy() for v in V])
    nu, dt = 0.1, 1e-2
    o.bdf1_chff(nu, dt, V)
    n = 16
    D2 = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1
        D2[:, k] = np.concatenate([C.chebder(e, 2), [0, 0]])
    A = nu * D2 - np.eye(n) / dt
    A[n - 2] = 1.0                           # u(1) = 0
    A[n - 1] = (-1.0) ** np.arange(n)        # u(-1) = 0
    for c in (1, 2):
        rhs = NL[c][:, 0, 0] - Vn[c][:, 0, 0] / dt
        rhs[n - 2:] = 0
        u = np.linalg.solve(A, rhs)
        assert np.abs(V[c][:, 0, 0] - u).max() <= 1e-12 * np.abs(u).max(), c
