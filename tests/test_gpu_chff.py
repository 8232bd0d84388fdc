"""The Chebyshev (ChFF) path on the GPU against the FP64 oracle (pinned in
test_oracle_chff.py): transforms, x-derivative, Helmholtz solves, the
nonlinear term and BDF1 steps with the influence-matrix correction."""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2, rel_l2_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402
from tarang_amd._lib import TarangGpuError  # noqa: E402

KF = (1.0, 0.5, 2.0)


def rnd(shape, seed, damp=True):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    if damp:
        A *= np.exp(-np.arange(shape[0]) / 4.0)[:, None, None]
    A[:, :, 0] = 0.5 * (A[:, :, 0] + np.conj(A[:, (-np.arange(shape[1])) % shape[1], 0]))
    return A


@pytest.mark.parametrize("N", [(32, 16, 16), (64, 8, 32)])
def test_chff_transforms_derivative_helmholtz(N):
    o = to.Oracle(N, kfactor=KF, dealias=to.DEALIAS_NONE, strong_reality=False)
    b = api.ChFF_Basis(N, kfactor=KF)
    r = np.zeros(o.rshape)
    r[:, :, :N[2]] = np.random.default_rng(1).standard_normal((N[0], N[1], N[2]))
    A = b.complex_array()
    b.Forward_transform(dev(r), A)
    ef = rel_l2(host(A), o.chff_forward(r))
    S = rnd(o.cshape, 2)
    Sd = dev(S)
    R = b.real_array()
    b.Inverse_transform(Sd, R)
    ei = rel_l2(host(R)[:, :, :N[2]], o.chff_inverse(S)[:, :, :N[2]])
    assert np.array_equal(host(Sd), S)
    B = b.complex_array()
    b.Xderiv(Sd, B)
    ed = rel_l2(host(B), o.cheb_xderiv(S))
    U = b.complex_array()
    b.Helmholtz(Sd, U, 3.5, 0.25 + 0.5j, -1.0j)
    eh = rel_l2(host(U), o.helmholtz(S, 3.5, 0.25 + 0.5j, -1.0j))
    log(f"ChFF N={N}: forward {ef:.2e} inverse {ei:.2e} xderiv {ed:.2e} helmholtz {eh:.2e}")
    assert ef < 1e-13 and ei < 1e-13 and ed < 1e-13 and eh < 1e-11


def test_chff_nonlinear_term():
    N = (32, 16, 16)
    o = to.Oracle(N, kfactor=KF, dealias=to.DEALIAS_NONE, strong_reality=False)
    V = [0.1 * rnd(o.cshape, 10 + c) for c in range(3)]
    b = api.ChFF_Basis(N, kfactor=KF)
    U = api.FluidVF(b, 0.05)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress_BDF1(U, 1e-3)
    ta.Compute_nlin()
    ref = o.compute_nlin_chff([v.copy() for v in V])
    e = rel_l2_vec([host(n) for n in U.nlin], ref)
    log(f"ChFF nonlinear term: rel L2 {e:.2e}")
    assert e < 1e-12


def _divfree_noslip(o, seed, amp=1e-2):
    """Divergence-free no-slip field exact in the retained modes (as in the oracle test)."""
    nx, ny, nzc = o.cshape
    rng = np.random.default_rng(seed)
    KY = np.fft.fftfreq(ny, 1 / ny) * o.kfactor[1]
    KY[ny // 2] = (ny // 2) * o.kfactor[1]
    w = np.array([0.5, 0.0, -0.5])
    V = [np.zeros(o.cshape, complex) for _ in range(3)]
    for ly in range(ny):
        for lz in range(nzc):
            ky, kz = KY[ly], lz * o.kfactor[2]
            if ly == ny // 2 or lz == nzc - 1 or (ly == 0 and lz == 0):
                continue
            K2 = ky * ky + kz * kz
            h = rng.standard_normal(7) + 1j * rng.standard_normal(7)
            q = rng.standard_normal(7) + 1j * rng.standard_normal(7)
            phi = C.chebmul(C.chebmul(w, w), h) / o.kfactor[0]
            chi = C.chebmul(w, q)
            dphi = C.chebder(phi) * o.kfactor[0]
            pad = lambda c: np.concatenate([c, np.zeros(nx - len(c))])
            V[0][:, ly, lz] = amp * pad(phi)
            V[1][:, ly, lz] = amp * (pad(1j * ky * dphi / K2) + pad(1j * kz * chi))
            V[2][:, ly, lz] = amp * (pad(1j * kz * dphi / K2) - pad(1j * ky * chi))
    for v in V:
        v[:, :, 0] = 0.5 * (v[:, :, 0] + np.conj(v[:, (-np.arange(ny)) % ny, 0]))
    return V


@pytest.mark.parametrize("N", [(32, 8, 16), (64, 16, 16)])
def test_bdf1_steps(N):
    o = to.Oracle(N, kfactor=KF, dealias=to.DEALIAS_NONE, strong_reality=False)
    V = _divfree_noslip(o, 20)
    nu, dt = 0.05, 2e-3
    b = api.ChFF_Basis(N, kfactor=KF)
    U = api.FluidVF(b, nu)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress_BDF1(U, dt)
    for s in range(3):
        P = o.bdf1_chff(nu, dt, V)
        ta.Time_advance_step(1)
        e = rel_l2_vec([host(v) for v in U.cvf], V)
        ep = rel_l2(host(ta.P), P)
        log(f"BDF1 N={N} step {s + 1}: velocity rel L2 {e:.2e} pressure {ep:.2e}")
        assert e < 1e-11 and ep < 1e-10
    # two steps in one call equal two single calls
    U2 = api.FluidVF(b, nu)
    U2.set_field([dev(v) for v in V])
    ta2 = api.Time_advance_incompress_BDF1(U2, dt)
    ta2.Time_advance_step(2)
    for _ in range(2):
        o.bdf1_chff(nu, dt, V)
    assert rel_l2_vec([host(v) for v in U2.cvf], V) < 1e-11


def test_chff_guards():
    b = api.ChFF_Basis((16, 8, 16))
    U = api.FluidVF(b, 0.1)
    with pytest.raises(TarangGpuError):
        api.Nlin_incompress.Compute_nlin(U)            # FFF-only
    with pytest.raises(TarangGpuError):
        b.Get_total_energy(U.cvf[0])                    # not defined on ChFF
    with pytest.raises(TarangGpuError):
        api.FFF_Basis.Forward_transform(b, b.real_array(), b.complex_array())   # Fourier code on ChFF
    f = api.FFF_Basis((16, 16, 16))
    Uf = api.FluidVF(f, 0.1)
    with pytest.raises(TarangGpuError):
        api.Time_advance_incompress_BDF1(Uf, 1e-3).Time_advance_step(1)
