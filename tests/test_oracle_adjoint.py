"""Pins of the oracle's tangent and adjoint restatement (CPU only).

The adjoint has no counterpart in the reference; these checks fix it by
mathematics alone: the tangent of the quadratic right-hand side against
central differences (exact up to round-off for a quadratic map), the
dot-product identity <J v, w> = <v, J^T w>, and the RK4-step and multi-step
gradients against central differences along seeded solenoidal directions.
"""
import numpy as np
import pytest

import tg_oracle as to


def dot(A, B):
    return sum(np.vdot(a, b).real for a, b in zip(A, B))


def rnd(o, rng, scale=0.1):
    return [(rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)).astype(np.complex128) * scale
            for _ in range(3)]


def rhs(o, V):
    V = [v.copy() for v in V]
    N = o.compute_nlin(V)
    o.compute_rhs(N)
    return N


def rk4(o, fl, V, n=1):
    V = [v.copy() for v in V]
    for _ in range(n):
        o.rk4_step(fl, V)
    return V


def shell_loss(o, V, r_lo, r_hi):
    kx = np.fft.fftfreq(o.nx, 1.0 / o.nx)
    ky = np.fft.fftfreq(o.ny, 1.0 / o.ny)
    kz = np.arange(o.nz // 2 + 1)
    K = np.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    w = np.ones(o.cshape)
    w[:, :, 0] = 0.5
    m = (K >= r_lo) & (K < r_hi)
    return sum((w * m * np.abs(A) ** 2).sum() for A in V)


@pytest.mark.parametrize("dealias", [to.DEALIAS_REF_CORNER, to.DEALIAS_TWOTHIRDS, to.DEALIAS_NONE])
def test_rhs_tangent_vs_central_difference(dealias):
    o = to.Oracle((16, 8, 16), dealias=dealias, strong_reality=False)
    rng = np.random.default_rng(5)
    V = o.init_field(kind=0, k0=3.0, seed=3)
    dV = rnd(o, rng)
    J = o.rhs_jvp(V, dV)
    h = 1e-3
    fp = rhs(o, [v + h * d for v, d in zip(V, dV)])
    fm = rhs(o, [v - h * d for v, d in zip(V, dV)])
    fd = [(a - b) / (2 * h) for a, b in zip(fp, fm)]
    err = np.sqrt(dot([a - b for a, b in zip(J, fd)], [a - b for a, b in zip(J, fd)]) / dot(fd, fd))
    assert err < 1e-12


@pytest.mark.parametrize("dealias", [to.DEALIAS_REF_CORNER, to.DEALIAS_TWOTHIRDS])
def test_rhs_adjoint_dot_product(dealias):
    o = to.Oracle((16, 16, 8), dealias=dealias, strong_reality=False)
    rng = np.random.default_rng(7)
    # non-Hermitian, non-solenoidal inputs: the identity holds on the whole
    # real vector space of stored coefficients
    V, v, w = rnd(o, rng, 1.0), rnd(o, rng), rnd(o, rng)
    lhs = dot(w, o.rhs_jvp(V, v))
    rhs_ = dot(o.rhs_vjp(V, w), v)
    assert abs(lhs - rhs_) <= 1e-13 * abs(lhs)


def test_dot_product_test_detects_a_dropped_weight():
    # the same identity with the kz weight c left out of the transpose fails:
    # the test above can tell a correct adjoint from a plausible wrong one
    o = to.Oracle((8, 8, 8), dealias=to.DEALIAS_NONE, strong_reality=False)
    rng = np.random.default_rng(9)
    V, v, w = rnd(o, rng, 1.0), rnd(o, rng), rnd(o, rng)
    good = o.rhs_vjp(V, w)
    c = np.full(o.cshape, 2.0)
    c[:, :, 0] = c[:, :, -1] = 1.0
    wrong = [g / c for g in good]
    lhs = dot(w, o.rhs_jvp(V, v))
    assert abs(lhs - dot(good, v)) <= 1e-13 * abs(lhs)
    assert abs(lhs - dot(wrong, v)) > 1e-3 * abs(lhs)


@pytest.mark.parametrize("hyper", [(0.0, 0), (2e-4, 4), (1e-5, 6)])
def test_rk4_step_adjoint_vs_central_difference(hyper):
    """Without and with hyperviscosity (nu_h, q)."""
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    fl = to.Fluid(0.01, 0.01, 0, 0.0, 0.0, 0.0, *hyper)
    rng = np.random.default_rng(11)
    V = o.init_field(kind=0, k0=3.0, seed=3)
    L, dV = rnd(o, rng), rnd(o, rng)
    g = o.rk4_step_vjp(fl, V, L)
    h = 3e-6
    fd = dot(L, [(a - b) / (2 * h) for a, b in zip(rk4(o, fl, [v + h * d for v, d in zip(V, dV)]),
                                                    rk4(o, fl, [v - h * d for v, d in zip(V, dV)]))])
    assert abs(fd - dot(g, dV)) <= 1e-9 * abs(fd)


def test_shell_energy_gradient_vs_central_difference():
    o = to.Oracle(16, dealias=to.DEALIAS_TWOTHIRDS, strong_reality=False)
    fl = to.Fluid(0.01, 0.01, 0, 0.0, 0.0, 0.0)
    V = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=7)
    nsteps, r_lo, r_hi = 3, 2.0, 5.0
    loss, grad = o.adjoint_shell_energy(fl, V, nsteps, r_lo, r_hi)
    assert abs(loss - shell_loss(o, rk4(o, fl, V, nsteps), r_lo, r_hi)) <= 1e-14 * loss
    eps = 1e-6
    for seed in range(8, 12):
        d = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=seed)  # solenoidal direction
        fp = shell_loss(o, rk4(o, fl, [v + eps * x for v, x in zip(V, d)], nsteps), r_lo, r_hi)
        fm = shell_loss(o, rk4(o, fl, [v - eps * x for v, x in zip(V, d)], nsteps), r_lo, r_hi)
        fd = (fp - fm) / (2 * eps)
        ad = dot(grad, d)
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), 1e-3 * abs(loss)), (seed, fd, ad)
