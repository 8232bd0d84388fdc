"""Failure behaviour of the time advance on the GPU: a non-finite state is
reported (TG_ERR_NAN) by the NS, MHD and RBC steps; a continuation of a step
chain is not taken across a change of the state or of the forcing band.
The non-finite values are planted from the host: the kernels only compute
with them (no out-of-range access follows from a NaN)."""
import ctypes

import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402
from tarang_amd._lib import ERR_NAN, ERR_STATE, STEP_CONTINUE, TarangGpuError, TgFluid  # noqa: E402


def _plant_nan(t):
    t[1, 2, 3] = complex(float("nan"), 0.0)


@pytest.mark.parametrize("prec", [api.PREC_FP64, api.PREC_FP32])
def test_ns_step_reports_nan(prec):
    N = (16, 16, 16)
    o = to.Oracle(N)
    V = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=1)
    b = api.FFF_Basis(N, precision=prec)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    ta = api.Time_advance_incompress(U, 1e-3)
    ta.Time_advance_step(2)          # finite: no error
    _plant_nan(U.cvf[1])
    with pytest.raises(TarangGpuError) as ei:
        ta.Time_advance_step(1)
    assert ei.value.code == ERR_NAN
    # a fresh finite state steps again without error on the same plan
    U.set_field([dev(v) for v in V])
    ta.Time_advance_step(1)


def test_mhd_and_rbc_steps_report_nan():
    N = (16, 16, 16)
    o = to.Oracle(N)
    V = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=2)
    W = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.2, seed=3)
    b = api.FFF_Basis(N)
    U, B = api.FluidVF(b, 0.01), api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    B.set_field([dev(w) for w in W])
    _plant_nan(B.cvf[0])
    with pytest.raises(TarangGpuError) as ei:
        api.Time_advance_incompress_MHD(U, B, 1e-3).Time_advance_step(1)
    assert ei.value.code == ERR_NAN
    kf = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))
    s = api.SFF_Basis(N, kfactor=kf, dealias=api.DEALIAS_REF_CORNER, reality=2)
    Us, Ts = api.FluidVF(s), api.FluidSF(s)
    rng = np.random.default_rng(4)
    shape = s.shape_complex_array
    for c in range(3):
        Us.cvf[c].copy_(dev(1e-3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))))
    Ts.set_field(dev(1e-3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))))
    ta = api.Time_advance_incompress_RBC(Us, Ts, 1e6, 1.0, 1e-4)
    ta.Time_advance_step(1)
    _plant_nan(Ts.F)
    with pytest.raises(TarangGpuError) as ei:
        ta.Time_advance_step(1)
    assert ei.value.code == ERR_NAN


def test_adjoint_backward_without_forward_is_a_state_error():
    b = api.FFF_Basis((16, 16, 16))
    U = api.FluidVF(b, 0.01)
    adj = api.Adjoint_incompress(U, 1e-3, checkpoint_every=2)
    lam = [b.complex_array() for _ in range(3)]
    with pytest.raises(TarangGpuError) as ei:
        adj.Backward(lam)
    assert ei.value.code == ERR_STATE


def test_chain_not_taken_after_set_field():
    N = (16, 16, 16)
    o = to.Oracle(N)
    V1 = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=5)
    V2 = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=6)
    b = api.FFF_Basis(N)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V1])
    ta = api.Time_advance_incompress(U, 1e-3)
    ta.Time_advance_step(1)
    U.set_field([dev(v) for v in V2])   # new state: the chain must restart
    ta.Time_advance_step(1)
    o.rk4_step(to.Fluid(0.01, 1e-3, 0, 0, 0, 0), V2)
    e = rel_l2_vec([host(v) for v in U.cvf], V2)
    log(f"chain after set_field: rel L2 {e:.3e}")
    assert e < 1e-12


def test_chain_not_taken_across_a_forcing_change():
    """An unforced call followed by a forced call flagged as a continuation:
    the band slots of the forced step are not prepared by the unforced call,
    so the library must restart the chain (the result equals the oracle's
    forced step)."""
    N = (16, 16, 16)
    o = to.Oracle(N)
    V = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=7)
    b = api.FFF_Basis(N)
    U = api.FluidVF(b, 0.01)
    U.set_field([dev(v) for v in V])
    f0 = TgFluid(0.01, 1e-3, 0, 0.0, 0.0, 0.0)
    f1 = TgFluid(0.01, 1e-3, 1, 0.1, 1.0, 3.0)
    for fl, flags in ((f0, 0), (f1, STEP_CONTINUE)):
        rc = b.L.tg_rk4_step(b.plan, U.ptrs(), ctypes.byref(fl), 1, flags)
        assert rc == 0
    o.rk4_step(to.Fluid(0.01, 1e-3, 0, 0, 0, 0), V)
    o.rk4_step(to.Fluid(0.01, 1e-3, 1, 0.1, 1.0, 3.0), V)
    e = rel_l2_vec([host(v) for v in U.cvf], V)
    log(f"chain across forcing change: rel L2 {e:.3e}")
    assert e < 1e-12
