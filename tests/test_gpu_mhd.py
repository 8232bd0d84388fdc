"""MHD (Elsasser) path on the GPU against the FP64 CPU oracle."""
import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2, rel_l2_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402

FP64_TOL = 1e-12


def _fields(o, s1, s2):
    V = o.init_field(kind=1, kmin=1.0, kmax=o.nx / 3, energy=0.5, seed=s1)
    W = o.init_field(kind=1, kmin=1.0, kmax=o.nx / 3, energy=0.25, seed=s2)
    return V, W


@pytest.mark.parametrize("dealias", [to.DEALIAS_REF_CORNER, to.DEALIAS_TWOTHIRDS])
def test_mhd_nlin(dealias):
    N = (32, 16, 32)
    o = to.Oracle(N, dealias=dealias)
    V, W = _fields(o, 21, 22)
    NU, NW = o.compute_nlin_mhd([v.copy() for v in V], [w.copy() for w in W])
    b = api.FFF_Basis(N, dealias=dealias)
    U, B = api.FluidVF(b), api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    B.set_field([dev(w) for w in W])
    api.Nlin_incompress.Compute_nlin(U, B, mhd=True)
    eu = rel_l2_vec([host(n) for n in U.nlin], NU)
    eb = rel_l2_vec([host(n) for n in B.nlin], NW)
    log(f"test_mhd_nlin dealias={dealias}: N_u {eu:.3e} N_b {eb:.3e}")
    assert eu < FP64_TOL and eb < FP64_TOL


@pytest.mark.parametrize("prec,tol", [(api.PREC_FP64, FP64_TOL), (api.PREC_FP32, 2e-5)])
def test_mhd_rk4_and_diagnostics(prec, tol):
    N = (32, 32, 32)
    o = to.Oracle(N)
    V, W = _fields(o, 23, 24)
    nu, eta, dt = 5e-3, 3e-3, 2e-3
    b = api.FFF_Basis(N, precision=prec)
    U, B = api.FluidVF(b, nu), api.FluidVF(b, eta)
    U.set_field([dev(v) for v in V])
    B.set_field([dev(w) for w in W])
    ta = api.Time_advance_incompress_MHD(U, B, dt)
    fl = to.Fluid(nu, dt, 0, 0, 0, 0)
    for s in range(3):
        o.rk4_step_mhd(fl, eta, V, W)
        ta.Time_advance_step(1)
        e = rel_l2_vec([host(v) for v in U.cvf] + [host(w) for w in B.cvf], V + W)
        log(f"  mhd step {s + 1} prec={prec}: rel L2 {e:.3e}")
        assert e < tol
    hc = api.Correlation.Get_cross_helicity(U, B)
    hc_ref = sum(o.cross_energy(V[c], W[c]) for c in range(3))
    assert abs(hc - hc_ref) <= max(tol, 1e-13) * abs(hc_ref) * 10
    if prec == api.PREC_FP64:
        radii = np.array([0.0, 2.0, 4.0, 8.0, 12.0, np.inf])
        u2b_ref, b2b_ref = o.flux_mhd([v.copy() for v in V], [w.copy() for w in W], radii)
        et = api.EnergyTr(b, radii, radii)
        u2b, b2b = et.Compute_flux_MHD(U, B)
        log(f"  mhd flux u2b rel {rel_l2(u2b, u2b_ref):.3e} b2b rel {rel_l2(b2b, b2b_ref):.3e}")
        assert rel_l2(u2b, u2b_ref) < 1e-11 and rel_l2(b2b, b2b_ref) < 1e-11


@pytest.mark.parametrize("strong", [True, False])
@pytest.mark.parametrize("prec,tol", [(api.PREC_FP64, FP64_TOL), (api.PREC_FP32, 2e-5)])
def test_mhd_steps_in_one_call(strong, prec, tol):
    """Three steps in one call: the stage-4 x pass leaves the next step's
    Elsasser input, and under the strong reality condition its kz = 0 and
    Nz/2 planes are re-formed after the condition.  The initial kz = 0 planes
    are perturbed off Hermitian symmetry so that the condition changes them
    after the first step."""
    N = (32, 16, 32)
    log(f"test_mhd_steps_in_one_call strong={strong} prec={prec}")
    o = to.Oracle(N, strong_reality=strong)
    V, W = _fields(o, 25, 26)
    rng = np.random.default_rng(27)
    for F in (V, W):
        for c in range(3):
            F[c][:, :, 0] += 1e-3 * (rng.standard_normal(F[c].shape[:2]) + 1j * rng.standard_normal(F[c].shape[:2]))
    nu, eta, dt = 5e-3, 3e-3, 2e-3
    b = api.FFF_Basis(N, precision=prec, strong_reality=strong)
    U, B = api.FluidVF(b, nu), api.FluidVF(b, eta)
    U.set_field([dev(v) for v in V])
    B.set_field([dev(w) for w in W])
    fl = to.Fluid(nu, dt, 0, 0, 0, 0)
    for _ in range(3):
        o.rk4_step_mhd(fl, eta, V, W)
    api.Time_advance_incompress_MHD(U, B, dt).Time_advance_step(3)
    e = rel_l2_vec([host(v) for v in U.cvf] + [host(w) for w in B.cvf], V + W)
    log(f"  after 3 steps: rel L2 {e:.3e}")
    assert e < tol
