"""Pins of the oracle's MHD time step, MHD fluxes, cross energy and seeded
initial spectra by checks that do not restate their formulas (CPU only):

* or_rk4_step_mhd: with b = 0 it is the Navier-Stokes step; with
  nu = eta = 0 and the 2/3 rule, total energy and cross helicity of the
  stepped state projected on the 2/3 box change by O(dt^5) or better per
  step (measured order);
* or_flux_mhd: u< -> b> and b< -> b> fluxes against a real-space evaluation
  in advective form, <b> . (b.grad) u<> and -<b> . (u.grad) b<>, on a
  band-limited field the grid resolves without aliasing; zero through a
  sphere enclosing every mode;
* or_cross_energy: Parseval against <a b> / 2 in real space;
* or_init_field: the per-mode amplitudes follow the prescribed spectrum
  (exactly for random phases, shell-averaged for Gaussian amplitudes).
"""
import numpy as np

import tg_oracle as to


def kgrid(o):
    kx = np.fft.fftfreq(o.nx, 1.0 / o.nx)
    ky = np.fft.fftfreq(o.ny, 1.0 / o.ny)
    kx[o.nx // 2] = o.nx // 2
    ky[o.ny // 2] = o.ny // 2
    kz = np.arange(o.nz // 2 + 1)
    return np.meshgrid(kx, ky, kz, indexing="ij")


def to_real(o, A):
    """Real field of a half spectrum (forward transform scaled by 1/N^3)."""
    return np.fft.irfftn(A, s=(o.nx, o.ny, o.nz), axes=(0, 1, 2)) * (o.nx * o.ny * o.nz)


def band_field(o, kmax, seed, energy):
    V = o.init_field(kind=1, kmin=1.0, kmax=kmax, energy=energy, seed=seed)
    return V


def test_mhd_step_with_zero_b_is_ns_step():
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    V = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=3)
    W = [np.zeros_like(v) for v in V]
    Vn = [v.copy() for v in V]
    fl = to.Fluid(0.02, 5e-3, 0, 0.0, 0.0, 0.0)
    o.rk4_step_mhd(fl, 0.01, V, W)
    o.rk4_step(fl, Vn)
    scale = max(np.abs(v).max() for v in Vn)
    assert max(np.abs(a - b).max() for a, b in zip(V, Vn)) <= 1e-15 * scale * 10
    assert all(np.abs(w).max() == 0.0 for w in W)


def _invariants(o, V, W):
    E = sum(o.energy(v) + o.energy(w) for v, w in zip(V, W))
    Hc = sum(o.cross_energy(v, w) for v, w in zip(V, W))
    return E, Hc


def test_mhd_step_conserves_invariants_to_fifth_order():
    """nu = eta = 0, 2/3 rule.  The step adds the nonlinear term's modes
    outside the 2/3 box to the state (the reference dealiases the state at
    the next nonlinear evaluation, not the nonlinear term), which changes the
    energy at O(dt^2); the projection of the stepped state onto the box is the
    Galerkin RK4 step, whose invariants change at O(dt^5) or better."""
    o = to.Oracle(16, dealias=to.DEALIAS_TWOTHIRDS, strong_reality=True)
    V0 = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=5)
    W0 = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.3, seed=6)
    for a in V0 + W0:
        o.dealias(a)
    E0, H0 = _invariants(o, V0, W0)
    dE, dH = [], []
    for dt in (0.02, 0.01):
        V, W = [v.copy() for v in V0], [w.copy() for w in W0]
        o.rk4_step_mhd(to.Fluid(0.0, dt, 0, 0.0, 0.0, 0.0), 0.0, V, W)
        for a in V + W:
            o.dealias(a)
        E, H = _invariants(o, V, W)
        dE.append(abs(E - E0))
        dH.append(abs(H - H0))
        # the state changed by O(dt), not by round-off
        assert max(np.abs(a - b).max() for a, b in zip(V, V0)) > 1e-3 * dt
    order_E = np.log2(dE[0] / dE[1])
    order_H = np.log2(dH[0] / dH[1])
    assert dE[1] > 1e-13 and dH[1] > 1e-14   # measurable above round-off
    assert order_E > 4.6, order_E
    assert order_H > 4.6, order_H
    # a wrong stage weight (e.g. 1/3 for 1/6) would leave a first-order change
    assert dE[1] < 1e-8 * E0 and dH[1] < 1e-8 * abs(H0)


def test_cross_energy_parseval():
    o = to.Oracle((16, 8, 16), dealias=to.DEALIAS_NONE)
    V = o.init_field(kind=1, kmin=1.0, kmax=3.0, energy=0.5, seed=7)
    W = o.init_field(kind=1, kmin=1.0, kmax=3.0, energy=0.2, seed=8)
    for c in range(3):
        a, b = to_real(o, V[c]), to_real(o, W[c])
        assert abs(o.cross_energy(V[c], W[c]) - 0.5 * (a * b).mean()) < 1e-15
        assert abs(o.cross_energy(V[c], V[c]) - o.energy(V[c])) < 1e-16


def _grad(o, A):
    KX, KY, KZ = kgrid(o)
    return [to_real(o, 1j * K * A) for K in (KX, KY, KZ)]


def _filter(o, A, r_lo, r_hi):
    KX, KY, KZ = kgrid(o)
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    return np.where((K > r_lo) & (K <= r_hi), A, 0)


def test_flux_mhd_against_real_space_advective_form():
    """Grid 24^3, fields excited for K <= 5: the products (K <= 10) are
    resolved without aliasing, so the divergence form the oracle evaluates
    equals the advective form evaluated here in real space."""
    o = to.Oracle(24, dealias=to.DEALIAS_NONE, strong_reality=True)
    V = band_field(o, 5.0, 11, 0.5)
    W = band_field(o, 5.0, 12, 0.25)
    radii = np.array([0.0, 1.5, 2.5, 3.5, 4.5, np.inf])
    u2b, b2b = o.flux_mhd([v.copy() for v in V], [w.copy() for w in W], radii)
    b_real = [to_real(o, w) for w in W]
    u_real = [to_real(o, v) for v in V]
    for s in range(1, len(radii)):
        R = radii[s]
        b_out = [to_real(o, _filter(o, w, R, np.inf)) for w in W]
        # u< -> b>:  < b> . (b . grad) u< >
        p = 0.0
        for i in range(3):
            g = _grad(o, _filter(o, V[i], -1.0, R))
            adv = sum(b_real[j] * g[j] for j in range(3))
            p += (b_out[i] * adv).mean()
        # b< -> b>:  -< b> . (u . grad) b< >
        q = 0.0
        for i in range(3):
            g = _grad(o, _filter(o, W[i], -1.0, R))
            adv = sum(u_real[j] * g[j] for j in range(3))
            q -= (b_out[i] * adv).mean()
        scale = np.abs(u2b).max() + np.abs(b2b).max()
        assert abs(u2b[s] - p) <= 1e-13 * scale, (s, u2b[s], p)
        assert abs(b2b[s] - q) <= 1e-13 * scale, (s, b2b[s], q)
    assert u2b[-1] == 0.0 and b2b[-1] == 0.0
    assert np.abs(u2b[1:-1]).max() > 1e-6 and np.abs(b2b[1:-1]).max() > 1e-6


def test_flux_mhd_brute_force_triad():
    """Three modes p + q = k on 8^3: u = u_p at p (and its conjugate), b =
    b_q at q and b_k at k; u< -> b> through R between |q| and |k| sums the
    transfer of every triad by hand."""
    o = to.Oracle(8, dealias=to.DEALIAS_NONE, strong_reality=True)
    p = np.array([1, 0, 1])
    q = np.array([0, 1, 1])
    k = p + q                    # (1, 1, 2), |k| = 2.45
    rng = np.random.default_rng(3)

    def solenoidal(kv):
        a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        kk = kv.astype(float)
        return a - kk * (kk @ a) / (kk @ kk)

    up, bq, bk = solenoidal(p), solenoidal(q), solenoidal(k)
    V = [o.czeros() for _ in range(3)]
    W = [o.czeros() for _ in range(3)]
    for c in range(3):
        V[c][p[0], p[1], p[2]] = up[c]
        W[c][q[0], q[1], q[2]] = bq[c]
        W[c][k[0], k[1], k[2]] = bk[c]
    radii = np.array([0.0, 2.0, np.inf])
    u2b, _ = o.flux_mhd([v.copy() for v in V], [w.copy() for w in W], radii)
    # transfer into b(k) from u(p) mediated by b(q): sum over the full
    # spectrum of Re[ b(k)^* . (i (b(q).p) u(p)) ], counted at k and -k
    T = 2 * np.real(np.conj(bk) @ (1j * (bq @ p) * up))
    assert abs(u2b[1] - T) <= 1e-14 * abs(T), (u2b[1], T)
    assert abs(T) > 1e-3


def test_init_field_random_phase_spectrum_exact():
    o = to.Oracle(32)
    kmin, kmax = 2.0, 12.0
    V = o.init_field(kind=1, k0=4.0, kmin=kmin, kmax=kmax, energy=0.5, seed=21)
    KX, KY, KZ = kgrid(o)
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    amp2 = sum(np.abs(v) ** 2 for v in V)
    nyq = (np.abs(KX) == o.nx // 2) | (np.abs(KY) == o.ny // 2) | (KZ == o.nz // 2)
    inside = (K >= kmin) & (K <= kmax) & ~nyq
    assert np.all(amp2[~inside] == 0.0)
    ratio = amp2[inside] * 2 * np.pi * K[inside] ** 2 / K[inside] ** (-5.0 / 3.0)
    assert ratio.max() - ratio.min() <= 1e-12 * ratio.mean()
    # shell-binned: E(K) sum over a shell of ~ 4 pi K^2 modes scales as K^(-5/3)
    sh = np.rint(K[inside]).astype(int)
    e_sh = np.bincount(sh, weights=amp2[inside] * np.where(KZ[inside] == 0, 0.5, 1.0) * 2)
    cnt = np.bincount(sh, weights=np.where(KZ[inside] == 0, 0.5, 1.0) * 2)
    ks = np.arange(len(e_sh))
    sel = (ks >= 4) & (ks <= 11)
    slope = np.polyfit(np.log(ks[sel]), np.log(e_sh[sel] / cnt[sel] * 4 * np.pi * ks[sel] ** 2), 1)[0]
    assert abs(slope + 5.0 / 3.0) < 0.05, slope


def test_init_field_gaussian_spectrum_shell_averages():
    o = to.Oracle(48)
    k0 = 4.0
    V = o.init_field(kind=0, k0=k0, kmin=0.0, kmax=100.0, energy=0.5, seed=22)
    KX, KY, KZ = kgrid(o)
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    amp2 = sum(np.abs(v) ** 2 for v in V)
    nyq = (np.abs(KX) == o.nx // 2) | (np.abs(KY) == o.ny // 2) | (KZ == o.nz // 2)
    ok = (K > 0) & ~nyq & (KZ > 0)     # kz = 0 half is a copy of the other half
    target = K[ok] ** 4 * np.exp(-2 * (K[ok] / k0) ** 2) / (2 * np.pi * K[ok] ** 2)
    r = amp2[ok] / target
    sh = np.rint(K[ok]).astype(int)
    mean = np.bincount(sh, weights=r) / np.maximum(np.bincount(sh), 1)
    n = np.bincount(sh)
    c = np.sum(r) / len(r)          # the normalisation constant
    for s in range(3, 15):
        # chi-square with 4 degrees of freedom per mode: relative sd 1/sqrt(n)
        assert abs(mean[s] / c - 1) < 4.0 / np.sqrt(n[s]), (s, mean[s] / c, n[s])
    assert abs(sum(o.energy(v) for v in V) - 0.5) < 1e-14
