"""Pins of the oracle's SFF basis and Rayleigh-Benard restatement (CPU only).

The reference's sin/cos transforms live in the absent fftk library; the
convention chosen (cell-centred samples, A = DCT-II/(2N) or shifted
DST-II/(2N)) is pinned by scipy.fft golden vectors, and everything built on
it by mathematics: exact single-mode transforms and derivatives, Parseval
against the reference's SFF energy formula, the nonlinear term against an
independent evaluation on the doubled periodic domain, the linear
Rayleigh-Benard growth rate of an eigenmode, and the Nusselt number against
its real-space definition.
"""
import os
import sys

import numpy as np
import pytest
from scipy.linalg import expm

import tg_oracle as to

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_sff_golden import sff_reference  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sff_golden.npz")
RBC_KF = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))  # SFF_pencil.cc:67-72


def oracle(N, dealias=to.DEALIAS_NONE, kf=RBC_KF):
    return to.Oracle(N, kfactor=kf, dealias=dealias, strong_reality=False)


def pad(r):
    nx, ny, nz = r.shape
    out = np.zeros((nx, ny, nz + 2))
    out[:, :, :nz] = r
    return out


def test_sff_transforms_match_scipy_golden():
    g = np.load(GOLDEN)
    for key in ["8x4x8", "16x8x4"]:
        r = g[f"real_{key}"]
        o = oracle(r.shape)
        for par, name in [(to.SIN, "sin"), (to.COS, "cos")]:
            A = o.sff_forward(par, pad(r))
            assert np.abs(A - g[f"{name}_{key}"]).max() < 1e-14


def test_single_modes_and_x_derivative():
    N = (16, 4, 4)
    o = oracle(N)
    x = (np.arange(N[0]) + 0.5) / N[0]  # L = 1 (kfactor pi)
    for m in (1, 3, 7):
        r = np.zeros(o.rshape)
        r[:, :, :N[2]] = 2 * np.sin(m * np.pi * x)[:, None, None]
        A = o.sff_forward(to.SIN, r)
        ref = o.czeros()
        ref[m, 0, 0] = 1.0
        assert np.abs(A - ref).max() < 1e-14
        # d/dx of a sine field: +Kx A in the cosine basis (SFF_pencil_tr.cc:84-88)
        dA = m * np.pi * A
        dr = o.sff_inverse(to.COS, dA)[:, 0, 0]
        assert np.abs(dr - 2 * m * np.pi * np.cos(m * np.pi * x)).max() < 1e-12
        r[:, :, :N[2]] = 2 * np.cos(m * np.pi * x)[:, None, None] + 1.0
        A = o.sff_forward(to.COS, r)
        ref[:] = 0
        ref[m, 0, 0] = 1.0
        ref[0, 0, 0] = 1.0
        assert np.abs(A - ref).max() < 1e-14


def _random_spec(o, rng, parity, mmax, kmax):
    A = o.czeros()
    nx, ny, nzc = o.cshape
    for lx in range(1 if parity == to.SIN else 0, mmax):
        for ly in range(ny):
            ky = ly if ly <= ny // 2 else ly - ny
            if abs(ky) >= kmax:
                continue
            for lz in range(kmax):
                A[lx, ly, lz] = rng.standard_normal() + 1j * rng.standard_normal()
    # a real field: Hermitian kz = 0 plane in ky for every x mode
    for lx in range(nx):
        for ly in range(1, ny // 2):
            A[lx, ny - ly, 0] = np.conj(A[lx, ly, 0])
        A[lx, 0, 0] = A[lx, 0, 0].real
    return A


def test_parseval_with_the_reference_energy_formula():
    N = (16, 8, 8)
    o = oracle(N)
    rng = np.random.default_rng(3)
    for par in (to.SIN, to.COS):
        A = _random_spec(o, rng, par, 12, 3)
        r = o.sff_inverse(par, A)[:, :, :N[2]]
        assert abs(o.sff_energy(A) - (r ** 2).sum() / (2 * np.prod(N))) < 1e-12 * o.sff_energy(A)
        assert np.abs(o.sff_forward(par, o.sff_inverse(par, A)) - A).max() < 1e-13


def _ext_deriv(f, parity_sign, axis, L, nx):
    """Spectral derivative of a field sampled on the doubled periodic x grid."""
    n = f.shape[axis]
    k = np.fft.fftfreq(n, 1.0 / n) * (2 * np.pi / L)
    F = np.fft.fft(f, axis=axis)
    shape = [1, 1, 1]
    shape[axis] = n
    return np.real(np.fft.ifft(1j * k.reshape(shape) * F, axis=axis))


def test_rbc_nonlinear_term_against_the_doubled_periodic_domain():
    # Alias-free fields (modes below a third of each axis): the pseudo-spectral
    # term is exact, so an independent evaluation must agree to round-off.
    N = (24, 12, 12)
    kf = (np.pi / 1.0, 2 * np.pi / 1.5, 2 * np.pi / 2.0)
    o = oracle(N, kf=kf)
    rng = np.random.default_rng(11)
    par = [to.SIN, to.COS, to.COS, to.SIN]
    S = [_random_spec(o, rng, p, N[0] // 3, 3) for p in par]  # products stay below the y, z Nyquist
    V, T = S[:3], S[3]
    Nu, NT = o.compute_nlin_rbc([v.copy() for v in V], T.copy())
    # real fields on the cell-centred grid, odd/even extension to 2 Nx
    fr = [o.sff_inverse(p, A)[:, :, :N[2]] for p, A in zip(par, S)]
    ext = [np.concatenate([f, (-1 if p == to.SIN else 1) * f[::-1]], axis=0) for f, p in zip(fr, par)]
    Ls = (2 * np.pi / kf[0] * 2 / 2, 2 * np.pi / kf[1], 2 * np.pi / kf[2])  # doubled x length = 2 L
    u, th = ext[:3], ext[3]

    def div(flux):  # sum_j d_j flux_j
        return sum(_ext_deriv(flux[j], 0, j, Ls[j], N[0]) for j in range(3))

    Nref = [div([u[j] * u[i] for j in range(3)]) for i in range(3)]
    NTref = div([u[j] * th for j in range(3)])
    outpar = [to.SIN, to.COS, to.COS]
    for i in range(3):
        ref = sff_reference(Nref[i][:N[0]], outpar[i])
        assert np.abs(Nu[i] - ref).max() < 1e-11 * np.abs(ref).max()
    ref = sff_reference(NTref[:N[0]], to.SIN)
    assert np.abs(NT - ref).max() < 1e-11 * np.abs(ref).max()


def _eigen_state(o, a, b, m=1, kyi=1):
    """u_x = 2a sin(m pi x) cos(ky y), u_y from continuity, theta = 2b sin(m pi x) cos(ky y)."""
    V = [o.czeros() for _ in range(3)]
    T = o.czeros()
    ny = o.cshape[1]
    Kx, Ky = m * o.g.kfactor[0], kyi * o.g.kfactor[1]
    for ly, s in [(kyi, 1), (ny - kyi, -1)]:
        V[0][m, ly, 0] = a / 2
        T[m, ly, 0] = b / 2
        # d_x u_x + d_y u_y = 0: +Kx Vx + i (s Ky) Vy = 0 (SFF divergence)
        V[1][m, ly, 0] = -Kx * (a / 2) / (1j * s * Ky)
    return V, T


def test_rbc_linear_growth_rate_of_an_eigenmode():
    N = (16, 8, 4)
    o = oracle(N, dealias=to.DEALIAS_TWOTHIRDS)
    Ra, Pr, dt, nsteps = 2000.0, 1.0, 2e-3, 50
    prm = to.Rbc(Ra, Pr, dt, 0)
    nu, kappa = np.sqrt(Pr / Ra), 1 / np.sqrt(Pr * Ra)
    Kx, Ky = np.pi, np.pi / np.sqrt(2)
    K2 = Kx ** 2 + Ky ** 2
    M = np.array([[-nu * K2, Ky ** 2 / K2], [1.0, -kappa * K2]])
    w, vecs = np.linalg.eig(M)
    ev = vecs[:, np.argmax(w.real)].real
    amp = 1e-9
    V, T = _eigen_state(o, amp * ev[0], amp * ev[1])
    for _ in range(nsteps):
        o.rk4_step_rbc(prm, V, T)
    ab = expm(M * dt * nsteps) @ (amp * ev)
    assert abs(V[0][1, 1, 0] * 2 - ab[0]) < 1e-9 * abs(ab[0])
    assert abs(T[1, 1, 0] * 2 - ab[1]) < 1e-9 * abs(ab[1])
    # above onset (Ra_c = 27 pi^4 / 4 = 657.5 for free-slip walls) the mode grows
    assert w.real.max() > 0


@pytest.mark.parametrize("dealias,order_ratio", [(to.DEALIAS_NONE, (14, 18)), (to.DEALIAS_TWOTHIRDS, (1.8, 2.6))])
def test_rk4_rbc_order_and_divergence_free(dealias, order_ratio):
    # Fourth order without dealiasing.  With the 2/3 rule the reference's
    # scheme is first order: stage 1 advances D V0 (Compute_nlin dealiases the
    # state in place, compute_nlin_SF.cc:79-82) while stages 2-3 and the final
    # update restart from the undealiased copy (RK_SF.cc:107,126,137,152), and
    # the buoyancy couples the dealiased modes of theta and u_x linearly.
    N = (16, 8, 8)
    o = oracle(N, dealias=dealias)
    rng = np.random.default_rng(5)
    V0, T0 = _eigen_state(o, 0.3, 0.2)
    T0 += 0.05 * _random_spec(o, rng, to.SIN, 4, 2)

    def run(dt, n):
        V, T = [v.copy() for v in V0], T0.copy()
        prm = to.Rbc(2000.0, 1.0, dt, 0)
        for _ in range(n):
            o.rk4_step_rbc(prm, V, T)
        return V, T

    ref = run(0.0025, 32)
    e = []
    for dt, n in [(0.02, 4), (0.01, 8)]:
        V, T = run(dt, n)
        e.append(np.sqrt(sum(np.abs(a - b).sum() ** 2 for a, b in zip(V + [T], ref[0] + [ref[1]]))))
    assert order_ratio[0] < e[0] / e[1] < order_ratio[1]
    # SFF divergence: +Kx Vx + i Ky Vy + i Kz Vz = 0
    V = ref[0]
    lx = np.arange(N[0])[:, None, None] * o.g.kfactor[0]
    ly = np.arange(N[1])
    ky = np.where(ly <= N[1] // 2, ly, ly - N[1])[None, :, None] * o.g.kfactor[1]  # ky = +Ny/2 at ly = Ny/2
    kz = np.arange(N[2] // 2 + 1)[None, None, :] * o.g.kfactor[2]
    d = lx * V[0] + 1j * ky * V[1] + 1j * kz * V[2]
    assert np.abs(d).max() < 1e-12


def test_nusselt_number():
    N = (16, 8, 8)
    o = oracle(N)
    prm = to.Rbc(1e4, 0.7, 1e-3, 2)
    zero = [o.czeros() for _ in range(3)]
    assert o.nusselt(prm, zero[0], o.czeros()) == 1.0   # conduction state
    V, T = _eigen_state(o, 0.3, 0.2)
    ux = o.sff_inverse(to.SIN, V[0])[:, :, :N[2]]
    th = o.sff_inverse(to.SIN, T)[:, :, :N[2]]
    # Nu = 1 + sqrt(Ra Pr) <u_x theta> (real-space mean over the sample grid)
    ref = 1 + np.sqrt(prm.Ra * prm.Pr) * (ux * th).mean()
    assert abs(o.nusselt(prm, V[0], T) - ref) < 1e-12 * ref


def test_sff_dealias_and_reality():
    N = (12, 12, 12)
    o = oracle(N, dealias=to.DEALIAS_REF_CORNER)
    A = np.ones(o.cshape, np.complex128)
    o.sff_dealias(A)
    # lx in [9, 11], ly in [5, 7], lz in [5, 6]: 3 * 3 * 2 zeros
    assert (A == 0).sum() == 18 and A[9:, 5:8, 5:].sum() == 0
    o2 = oracle(N, dealias=to.DEALIAS_TWOTHIRDS)
    A = np.ones(o2.cshape, np.complex128)
    o2.sff_dealias(A)
    ky = np.abs(np.fft.fftfreq(12, 1 / 12))
    keep = (np.arange(12)[:, None, None] <= 8) & (ky[None, :, None] <= 4) & (np.arange(7)[None, None, :] <= 4)
    assert np.array_equal(A != 0, keep)
    rng = np.random.default_rng(2)
    B = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    C = B.copy()
    o.sff_reality(C, 2)
    assert np.all(C[:, :, -1] == 0) and np.array_equal(C[:, :, :-1], B[:, :, :-1])
    C = B.copy()
    o.sff_reality(C, 1)
    for ly in range(1, 6):
        assert np.array_equal(C[:, 12 - ly, 0], np.conj(B[:, ly, 0]))
    assert np.array_equal(C[:, 0, 0], np.conj(B[:, 0, 0]))      # the reference conjugates ky = 0
    assert np.array_equal(C[:, 6, 0], B[:, 6, 0])               # ky = Ny/2 untouched
