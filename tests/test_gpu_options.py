"""Options of the reference's time advance beyond the plain configuration,
on the GPU against the FP64 oracle (whose forms are pinned in
test_oracle.py and test_oracle_adjoint.py): hyperviscosity (FluidVF
hyper_dissipation, Array_exp_ksqr with hyper_factor) in the RK4 step and in
its adjoint."""
import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402
from tarang_amd._lib import ERR_ARG, TarangGpuError  # noqa: E402

TOL = {api.PREC_FP64: 1e-12, api.PREC_FP32: 2e-5}


@pytest.mark.parametrize("hyper", [(1e-5, 4), (2e-9, 6)])
@pytest.mark.parametrize("prec", [api.PREC_FP64, api.PREC_FP32])
def test_hyperviscous_rk4_steps(hyper, prec):
    N = (32, 32, 32)
    nuh, q = hyper
    log(f"test_hyperviscous_rk4_steps nu_h={nuh} q={q} prec={prec}")
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=12.0, energy=0.5, seed=51)
    force = (1, 0.1, 1.0, 3.0)
    b = api.FFF_Basis(N, precision=prec, strong_reality=True)
    U = api.FluidVF(b, 0.005, hyper_dissipation_coefficient=nuh, hyper_dissipation_exponent=q)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, 2e-3, *force)
    fl = to.Fluid(0.005, 2e-3, *force, nuh, q)
    E0 = sum(o.energy(v) for v in V)
    for s in range(3):
        o.rk4_step(fl, V)
        ta.Time_advance_step(1)
        e = rel_l2_vec([host(v) for v in U.cvf], V)
        log(f"  step {s + 1}: rel L2 {e:.2e}")
        assert e < TOL[prec]
    # the hyperviscous run dissipates more than the plain one from the same field
    W = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=12.0, energy=0.5, seed=51)
    for _ in range(3):
        o.rk4_step(to.Fluid(0.005, 2e-3, *force), W)
    assert sum(o.energy(v) for v in V) < sum(o.energy(w) for w in W) < E0 + 1e-3


def test_hyperviscous_adjoint_gradient():
    N = (32, 32, 32)
    nu, dt, nuh, q = 0.01, 0.01, 1e-5, 4
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    fl = to.Fluid(nu, dt, 0, 0.0, 0.0, 0.0, nuh, q)
    V = o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=41)
    nsteps, r_lo, r_hi = 5, 2.0, 5.0
    loss_ref, grad_ref = o.adjoint_shell_energy(fl, V, nsteps, r_lo, r_hi)
    b = api.FFF_Basis(N, strong_reality=False)
    U = api.FluidVF(b, nu, hyper_dissipation_coefficient=nuh, hyper_dissipation_exponent=q)
    U.set_field([dev(v) for v in V])
    adj = api.Adjoint_incompress(U, dt, checkpoint_every=2)
    adj.Forward(nsteps)
    # E in r_lo <= K < r_hi as the oracle's loss: sum |V|^2 with weight 1/2 on kz = 0
    kx = np.fft.fftfreq(N[0], 1.0 / N[0])[:, None, None]
    ky = np.fft.fftfreq(N[1], 1.0 / N[1])[None, :, None]
    kz = np.arange(N[2] // 2 + 1)[None, None, :]
    K = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    w = np.where(kz == 0, 0.5, 1.0)
    m = torch.as_tensor(w * ((K >= r_lo) & (K < r_hi)), device="cuda")
    loss = float(sum((m * (v.real ** 2 + v.imag ** 2)).sum().item() for v in U.cvf))
    lam = [2.0 * m * v for v in U.cvf]
    adj.Backward(lam)
    eg = rel_l2_vec([host(x) for x in lam], grad_ref)
    log(f"test_hyperviscous_adjoint_gradient: loss {abs(loss - loss_ref) / loss_ref:.2e} grad {eg:.2e}")
    assert abs(loss - loss_ref) < 1e-12 * loss_ref and eg < 1e-11


def test_hyperviscosity_arguments():
    b = api.FFF_Basis((16, 16, 16))
    U = api.FluidVF(b, 0.01, hyper_dissipation_coefficient=1e-4, hyper_dissipation_exponent=1)
    with pytest.raises(TarangGpuError) as e:
        api.Time_advance_incompress(U, 1e-3).Time_advance_step(1)
    assert e.value.code == ERR_ARG
    U = api.FluidVF(b, 0.01, hyper_dissipation_coefficient=1e-4, hyper_dissipation_exponent=4)
    W = api.FluidVF(b, 0.01)
    with pytest.raises(TarangGpuError) as e:
        api.Time_advance_incompress_MHD(U, W, 1e-3).Time_advance_step(1)   # MHD takes no hyperviscosity
    assert e.value.code == ERR_ARG


@pytest.mark.parametrize("prec", [api.PREC_FP64, api.PREC_FP32])
def test_get_dt_fff(prec):
    """FluidVF::Get_dt (dt_option 1) of a seeded field against the oracle;
    the dt_fixed cap; a non-finite field is reported."""
    N = (32, 16, 48) if prec == api.PREC_FP64 else (32, 16, 64)
    kf = (1.0, 0.5, 2.0)
    o = to.Oracle(N, kfactor=kf, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=10.0, energy=0.5, seed=61)
    b = api.FFF_Basis(N, kfactor=kf, precision=prec, strong_reality=True)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    dt_ref, um_ref = o.get_dt([v.copy() for v in V], 0.5, 1.0)
    dt = U.Get_dt(0.5, 1.0)
    tol = 1e-12 if prec == api.PREC_FP64 else 1e-6
    log(f"test_get_dt_fff prec={prec}: dt {dt:.6e} ref {dt_ref:.6e} umax {U.umax} ref {tuple(um_ref)}")
    assert abs(dt - dt_ref) <= tol * dt_ref
    assert np.allclose(U.umax, um_ref, rtol=tol, atol=0)
    assert U.Get_dt(0.5, 0.5 * dt_ref) == 0.5 * dt_ref
    assert U.Get_dt(0.5, 0.1, dt_option=0) == 0.1
    bad = host(U.cvf[1])
    bad[2, 3, 4] = np.nan
    U.cvf[1].copy_(torch.as_tensor(bad))
    with pytest.raises(TarangGpuError) as e:
        U.Get_dt(0.5, 1.0)
    assert e.value.code == to_err_nan()


def to_err_nan():
    from tarang_amd._lib import ERR_NAN
    return ERR_NAN


def test_get_dt_sff():
    kf = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))
    N = (32, 16, 32)
    o = to.Oracle(N, kfactor=kf, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    rng = np.random.default_rng(62)
    V = [0.1 * np.exp(-np.arange(N[0]) / 4.0)[:, None, None] * (rng.standard_normal(o.cshape) + 1j *
                                                              rng.standard_normal(o.cshape)) for _ in range(3)]
    V[0][0] = 0
    b = api.SFF_Basis(N, kfactor=kf, dealias=api.DEALIAS_REF_CORNER, reality=2)
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    dt_ref, um_ref = o.get_dt([v.copy() for v in V], 0.3, 1.0, sff=True)
    dt = U.Get_dt(0.3, 1.0)
    log(f"test_get_dt_sff: dt {dt:.6e} ref {dt_ref:.6e}")
    assert abs(dt - dt_ref) <= 1e-12 * dt_ref
    assert np.allclose(U.umax, um_ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("prec", [api.PREC_FP64, api.PREC_FP32])
def test_random_band_forcing(prec):
    """Random band forcing (force_kind 2): the forced right-hand side and
    RK4 steps against the oracle; the phases are redrawn (seed + 1) once a
    random interval has passed."""
    N = (32, 32, 32)
    nu, dt, tau, seed = 0.01, 1e-3, 1e-3, 1234
    frc = dict(force_kind=2, force_eps=0.2, force_r_in=2.0, force_r_out=4.0, force_helicity=0.3, force_tau=tau,
               force_seed=seed)
    log(f"test_random_band_forcing prec={prec}")
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = o.init_field(kind=1, k0=4.0, kmin=1.0, kmax=10.0, energy=0.5, seed=71)

    def ofl(s):
        return to.Fluid(nu, dt, 2, 0.2, 2.0, 4.0, 0.0, 0, 0.3, tau, s)

    b = api.FFF_Basis(N, precision=prec, strong_reality=True)
    U = api.FluidVF(b, nu)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, dt, **frc)
    ta.Compute_force_TO_rhs()
    Vd = [v.copy() for v in V]
    for v in Vd:
        o.dealias(v)
    rhs_ref = o.compute_nlin([v.copy() for v in Vd])
    o.compute_rhs(rhs_ref, o.compute_force(ofl(seed), Vd))      # in place
    e_r = rel_l2_vec([host(n) for n in U.nlin], rhs_ref)
    log(f"  forced rhs rel L2 {e_r:.2e}")
    assert e_r < TOL[prec]
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, dt, **frc)
    for s in range(3):
        ta.Time_advance_step(1)
        o.rk4_step(ofl(seed + s), V)      # a new interval each step (tau = dt)
        e = rel_l2_vec([host(v) for v in U.cvf], V)
        log(f"  step {s + 1} (seed {ta.fluid.force_seed}): rel L2 {e:.2e}")
        assert ta.fluid.force_seed == seed + s and e < TOL[prec]
