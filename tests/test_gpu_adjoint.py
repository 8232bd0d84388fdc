"""Adjoint of the NS right-hand side and of Lawson RK4 on the GPU, against the
FP64 CPU adjoint of the oracle and central differences of the GPU forward map.

The adjoint has no counterpart in the reference.  The oracle's adjoint is
pinned by the dot-product identity and central differences on the CPU
(test_oracle_adjoint.py); here the GPU path is held to it at 1e-12 (FP64).
"""
import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402
from tarang_amd._lib import TarangGpuError  # noqa: E402

FP64_TOL = 1e-12


def dot(A, B):
    return sum(np.vdot(a, b).real for a, b in zip(A, B))


def rnd(shape, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * scale for _ in range(3)]


def shell_mask(N, r_lo, r_hi):
    nx, ny, nz = N
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    kz = np.arange(nz // 2 + 1)
    K = np.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    w = np.ones(K.shape)
    w[:, :, 0] = 0.5
    return w * ((K >= r_lo) & (K < r_hi))


def make(N, prec=api.PREC_FP64, dealias=api.DEALIAS_REF_CORNER, nu=0.01):
    b = api.FFF_Basis(N, precision=prec, dealias=dealias, strong_reality=False)
    return b, api.FluidVF(b, nu)


@pytest.mark.parametrize("dealias", [api.DEALIAS_REF_CORNER, api.DEALIAS_TWOTHIRDS])
@pytest.mark.parametrize("prec,tol", [(api.PREC_FP64, FP64_TOL), (api.PREC_FP32, 2e-5)])
def test_rhs_vjp_vs_oracle(dealias, prec, tol):
    N = (32, 16, 32)
    o = to.Oracle(N, dealias=dealias, strong_reality=False)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=31)
    Lam = rnd(o.cshape, 32)
    ref = o.rhs_vjp(V, Lam)
    b, U = make(N, prec, dealias)
    U.set_field([dev(v) for v in V])
    adj = api.Adjoint_incompress(U, 1e-3)
    out = adj.Rhs_vjp([dev(x) for x in Lam])
    got = [host(x) for x in out]
    e = rel_l2_vec(got, ref)
    # V is read only
    assert rel_l2_vec([host(v) for v in U.cvf], V) == 0.0
    # dot-product identity with the oracle's tangent
    v = rnd(o.cshape, 33)
    lhs = dot(Lam, o.rhs_jvp(V, v))
    rel_dot = abs(lhs - dot(got, v)) / abs(lhs)
    log(f"test_rhs_vjp dealias={dealias} prec={prec}: rel L2 {e:.3e}, dot identity {rel_dot:.3e}")
    assert e < tol and rel_dot < 10 * tol


def _gpu_gradient(N, V, nsteps, every, r_lo, r_hi, prec=api.PREC_FP64, dealias=api.DEALIAS_REF_CORNER,
                  nu=0.01, dt=0.01):
    b, U = make(N, prec, dealias, nu)
    U.set_field([dev(v) for v in V])
    adj = api.Adjoint_incompress(U, dt, checkpoint_every=every)
    adj.Forward(nsteps)
    Vn = [host(v) for v in U.cvf]
    m = torch.as_tensor(shell_mask(N, r_lo, r_hi), device="cuda")
    lam = [2.0 * m * v for v in U.cvf]
    loss = float(sum((m * (v.real ** 2 + v.imag ** 2)).sum().item() for v in U.cvf))
    adj.Backward(lam)
    return loss, Vn, [host(x) for x in lam]


def test_adjoint_gradient_vs_oracle_with_recompute():
    N = (32, 32, 32)
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    fl = to.Fluid(0.01, 0.01, 0, 0.0, 0.0, 0.0)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=41)
    nsteps, r_lo, r_hi = 5, 2.0, 5.0
    loss_ref, grad_ref = o.adjoint_shell_energy(fl, V, nsteps, r_lo, r_hi)
    Vn_ref = [v.copy() for v in V]
    for _ in range(nsteps):
        o.rk4_step(fl, Vn_ref)
    # checkpoints every 2 steps: segments [0,2), [2,4), [4,5); two recomputed
    loss, Vn, grad = _gpu_gradient(N, V, nsteps, 2, r_lo, r_hi)
    ev, eg = rel_l2_vec(Vn, Vn_ref), rel_l2_vec(grad, grad_ref)
    log(f"test_adjoint_gradient: V_n {ev:.3e} loss {abs(loss - loss_ref) / loss_ref:.3e} grad {eg:.3e}")
    assert ev < FP64_TOL and abs(loss - loss_ref) < 1e-12 * loss_ref and eg < 1e-11


def test_checkpoint_interval_does_not_change_the_gradient():
    N = (32, 32, 32)
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=42)
    grads = [_gpu_gradient(N, V, 6, every, 2.0, 6.0)[2] for every in (1, 4, 6, 50)]
    for g in grads[1:]:
        e = rel_l2_vec(g, grads[0])
        log(f"  checkpoint interval: rel diff {e:.3e}")
        assert e < 1e-14


def test_adjoint_vs_central_differences_of_the_gpu_forward():
    # 16 seeded solenoidal directions, eps = 1e-6, through the GPU forward map
    N = (32, 32, 32)
    nsteps, r_lo, r_hi, nu, dt = 10, 4.0, 8.0, 0.01, 0.005
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=7)
    loss, _, grad = _gpu_gradient(N, V, nsteps, 4, r_lo, r_hi, nu=nu, dt=dt)
    b, U = make(N, nu=nu)
    ta = api.Time_advance_incompress(U, dt)
    m = torch.as_tensor(shell_mask(N, r_lo, r_hi), device="cuda")

    def f(Vs):
        U.set_field([dev(v) for v in Vs])
        ta.Time_advance_step(nsteps, chain=False)
        return float(sum((m * (v.real ** 2 + v.imag ** 2)).sum().item() for v in U.cvf))

    eps = 1e-6
    worst = 0.0
    for seed in range(8, 24):
        d = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=seed)
        fd = (f([v + eps * x for v, x in zip(V, d)]) - f([v - eps * x for v, x in zip(V, d)])) / (2 * eps)
        ad = dot(grad, d)
        rel = abs(fd - ad) / max(abs(fd), 1e-3 * loss)
        worst = max(worst, rel)
    log(f"test_adjoint_fd: loss {loss:.6e}, worst relative FD mismatch over 16 directions {worst:.3e}")
    assert worst < 1e-6


def test_fp32_adjoint_gradient():
    N = (32, 32, 32)
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    fl = to.Fluid(0.01, 0.01, 0, 0.0, 0.0, 0.0)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=43)
    _, grad_ref = o.adjoint_shell_energy(fl, V, 4, 2.0, 5.0)
    _, _, grad = _gpu_gradient(N, V, 4, 2, 2.0, 5.0, prec=api.PREC_FP32)
    e = rel_l2_vec(grad, grad_ref)
    log(f"test_fp32_adjoint_gradient: rel L2 {e:.3e}")
    assert e < 1e-5


class _Rk4Map(torch.autograd.Function):
    """Test harness: nsteps of the GPU RK4 map as a torch autograd node whose
    backward is the library's checkpointed adjoint."""

    @staticmethod
    def forward(ctx, V, adj, nsteps):
        for c in range(3):
            adj.U.cvf[c].copy_(V[c])
        adj.Forward(nsteps)
        ctx.adj = adj
        return torch.stack([v.clone() for v in adj.U.cvf])

    @staticmethod
    def backward(ctx, g):
        lam = [g[c].contiguous().clone() for c in range(3)]
        ctx.adj.Backward(lam)
        return torch.stack(lam), None, None


def test_torch_autograd_harness():
    N = (32, 32, 32)
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    fl = to.Fluid(0.01, 0.01, 0, 0.0, 0.0, 0.0)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=44)
    _, grad_ref = o.adjoint_shell_energy(fl, V, 3, 2.0, 5.0)
    b, U = make(N)
    adj = api.Adjoint_incompress(U, 0.01, checkpoint_every=2)
    V0 = torch.stack([dev(v) for v in V]).requires_grad_(True)
    Vn = _Rk4Map.apply(V0, adj, 3)
    m = torch.as_tensor(shell_mask(N, 2.0, 5.0), device="cuda")
    loss = (m * (Vn.real ** 2 + Vn.imag ** 2)).sum()
    loss.backward()
    e = rel_l2_vec([host(V0.grad[c]) for c in range(3)], grad_ref)
    log(f"test_torch_autograd_harness: rel L2 {e:.3e}")
    assert e < 1e-11


def test_adjoint_argument_errors():
    N = (16, 16, 16)
    b, U = make(N)
    adj = api.Adjoint_incompress(U, 0.01)
    lam = [torch.zeros_like(U.cvf[0]) for _ in range(3)]
    with pytest.raises(TarangGpuError):
        adj.Backward(lam)  # nothing recorded
    forced = api.Adjoint_incompress(U, 0.01)
    forced.fluid.force_kind = 1
    with pytest.raises(TarangGpuError):
        forced.Forward(2)
    br = api.FFF_Basis(N)  # strong reality on
    Ur = api.FluidVF(br, 0.01)
    with pytest.raises(TarangGpuError):
        api.Adjoint_incompress(Ur, 0.01).Forward(1)
    adj.Forward(2)
    adj.Free()
    with pytest.raises(TarangGpuError):
        adj.Backward(lam)
