"""SFF basis and Rayleigh-Benard convection on the GPU against the FP64 CPU
oracle (tests/test_oracle_sff.py pins the oracle)."""
import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, host, log, rel_l2, rel_l2_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402
from tarang_amd._lib import TarangGpuError  # noqa: E402

FP64_TOL = 1e-12
RBC_KF = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))


def smooth_spec(shape, kf, seed, k0=6.0, sine=True):
    """Random coefficients with a Gaussian envelope in |K| (A_0 = 0 for sine)."""
    rng = np.random.default_rng(seed)
    nx, ny, nzc = shape
    lx = np.arange(nx)[:, None, None] * kf[0]
    ly = np.arange(ny)
    ky = np.where(ly <= ny // 2, ly, ly - ny)[None, :, None] * kf[1]
    kz = np.arange(nzc)[None, None, :] * kf[2]
    K = np.sqrt(lx ** 2 + ky ** 2 + kz ** 2)
    A = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * np.exp(-(K / (k0 * kf[0])) ** 2)
    if sine:
        A[0] = 0
    return A.astype(np.complex128)


def rbc_state(shape, kf, seed):
    V = [smooth_spec(shape, kf, seed + c, sine=(c == 0)) * 0.1 for c in range(3)]
    T = smooth_spec(shape, kf, seed + 3, sine=True) * 0.1
    return V, T


def make(N, prec=api.PREC_FP64, dealias=api.DEALIAS_REF_CORNER, reality=2, kf=RBC_KF):
    return api.SFF_Basis(N, kfactor=kf, precision=prec, dealias=dealias, reality=reality)


@pytest.mark.parametrize("N", [(16, 8, 16), (32, 16, 32), (64, 32, 16)])
@pytest.mark.parametrize("prec,tol", [(api.PREC_FP64, FP64_TOL), (api.PREC_FP32, 1e-5)])
def test_sff_transforms(N, prec, tol):
    o = to.Oracle(N, kfactor=RBC_KF, dealias=to.DEALIAS_NONE, strong_reality=False)
    b = make(N, prec, api.DEALIAS_NONE)
    rng = np.random.default_rng(1)
    r = np.zeros(o.rshape)
    r[:, :, :N[2]] = rng.standard_normal((N[0], N[1], N[2]))
    for sc, par in [(api.SC_SIN, to.SIN), (api.SC_COS, to.COS)]:
        ref = o.sff_forward(par, r)
        A = b.complex_array()
        b.Forward_transform(dev(r), A, sc)
        ef = rel_l2(host(A), ref)
        back = b.real_array()
        b.Inverse_transform(A, back, sc)
        rback = o.sff_inverse(par, host(A))
        eb = rel_l2(host(back)[:, :, :N[2]], rback[:, :, :N[2]])
        log(f"test_sff_transforms N={N} prec={prec} sincos={sc}: forward {ef:.3e} inverse {eb:.3e}")
        assert ef < tol and eb < tol


@pytest.mark.parametrize("dealias", [api.DEALIAS_REF_CORNER, api.DEALIAS_TWOTHIRDS])
def test_rbc_nonlinear_terms(dealias):
    N = (32, 16, 32)
    o = to.Oracle(N, kfactor=RBC_KF, dealias=dealias, strong_reality=False)
    V, T = rbc_state(o.cshape, RBC_KF, 10)
    Nref, NTref = o.compute_nlin_rbc([v.copy() for v in V], T.copy())
    b = make(N, dealias=dealias)
    U, S = api.FluidVF(b), api.FluidSF(b)
    U.set_field([dev(v) for v in V])
    S.set_field(dev(T))
    ta = api.Time_advance_incompress_RBC(U, S, 1e4, 1.0, 1e-3)
    ta.Compute_nlin()
    e = rel_l2_vec([host(n) for n in U.nlin], Nref)
    et = rel_l2(host(S.nlin), NTref)
    # the state is dealiased in place, as in the reference
    Vd = [v.copy() for v in V]
    for v in Vd:
        o.sff_dealias(v)
    ed = rel_l2_vec([host(v) for v in U.cvf], Vd)
    log(f"test_rbc_nonlinear_terms dealias={dealias}: N {e:.3e} N_T {et:.3e} state {ed:.1e}")
    assert e < FP64_TOL and et < FP64_TOL and ed == 0.0


def test_rbc_rhs():
    N = (32, 16, 16)
    o = to.Oracle(N, kfactor=RBC_KF, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    V, T = rbc_state(o.cshape, RBC_KF, 20)
    Vo, To = [v.copy() for v in V], T.copy()
    Fu1, FT = To.copy(), Vo[0].copy()
    Nref, NTref = o.compute_nlin_rbc(Vo, To)
    o.compute_rhs_rbc(Fu1, FT, Nref, NTref)
    b = make(N)
    U, S = api.FluidVF(b), api.FluidSF(b)
    U.set_field([dev(v) for v in V])
    S.set_field(dev(T))
    api.Time_advance_incompress_RBC(U, S, 1e4, 1.0, 1e-3).Compute_force_TO_rhs()
    e = rel_l2_vec([host(n) for n in U.nlin] + [host(S.nlin)], Nref + [NTref])
    log(f"test_rbc_rhs: {e:.3e}")
    assert e < FP64_TOL


@pytest.mark.parametrize("dealias,reality", [(api.DEALIAS_REF_CORNER, 2), (api.DEALIAS_TWOTHIRDS, 0),
                                             (api.DEALIAS_REF_CORNER, 1)])
@pytest.mark.parametrize("prec,tol", [(api.PREC_FP64, FP64_TOL), (api.PREC_FP32, 2e-5)])
def test_rbc_rk4(dealias, reality, prec, tol):
    N = (32, 16, 32)
    Ra, Pr, dt = 1e4, 0.7, 2e-3
    o = to.Oracle(N, kfactor=RBC_KF, dealias=dealias, strong_reality=False)
    V, T = rbc_state(o.cshape, RBC_KF, 30)
    b = make(N, prec, dealias, reality)
    U, S = api.FluidVF(b), api.FluidSF(b)
    U.set_field([dev(v) for v in V])
    S.set_field(dev(T))
    ta = api.Time_advance_incompress_RBC(U, S, Ra, Pr, dt)
    prm = to.Rbc(Ra, Pr, dt, reality)
    for s in range(3):
        o.rk4_step_rbc(prm, V, T)
        ta.Time_advance_step(1)
        e = rel_l2_vec([host(v) for v in U.cvf] + [host(S.F)], V + [T])
        log(f"  rbc step {s + 1} dealias={dealias} reality={reality} prec={prec}: rel L2 {e:.3e}")
        assert e < tol
    # energies with the SFF weights and the Nusselt number
    eg = [b.Get_total_energy(a) for a in U.cvf + [S.F]]
    eo = [o.sff_energy(a) for a in V + [T]]
    nu = api.Correlation.Get_Nusselt_no(U, S, Ra, Pr)
    nuo = o.nusselt(prm, V[0], T)
    log(f"  energies {eg} vs {eo}; Nu {nu:.15g} vs {nuo:.15g}")
    for a, c in zip(eg, eo):
        assert abs(a - c) <= max(tol, 1e-13) * abs(c) * 10
    assert abs(nu - nuo) <= max(tol, 1e-13) * abs(nuo) * 10


def test_rbc_multi_step_call_equals_single_steps():
    N = (16, 8, 16)
    V, T = rbc_state((16, 8, 9), RBC_KF, 40)
    out = []
    for split in (False, True):
        b = make(N)
        U, S = api.FluidVF(b), api.FluidSF(b)
        U.set_field([dev(v) for v in V])
        S.set_field(dev(T))
        ta = api.Time_advance_incompress_RBC(U, S, 1e4, 1.0, 1e-3)
        if split:
            for _ in range(4):
                ta.Time_advance_step(1)
        else:
            ta.Time_advance_step(4)
        out.append([host(v) for v in U.cvf] + [host(S.F)])
    assert rel_l2_vec(out[0], out[1]) == 0.0


def test_sff_reality_and_dealias():
    N = (16, 32, 16)
    o = to.Oracle(N, kfactor=RBC_KF, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    rng = np.random.default_rng(3)
    A = (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)).astype(np.complex128)
    b = make(N)
    for mode in (1, 2):
        ref = A.copy()
        o.sff_reality(ref, mode)
        d = dev(A)
        if mode == 1:
            b.Satisfy_strong_reality_condition_in_Array(d)
        else:
            b.Satisfy_weak_reality_condition_in_Array(d)
        assert np.array_equal(host(d), ref)
    ref = A.copy()
    o.sff_dealias(ref)
    d = dev(A)
    b.Dealias(d)
    assert np.array_equal(host(d), ref)


def test_basis_guards():
    b = make((16, 8, 16))
    U = api.FluidVF(b)
    with pytest.raises(TarangGpuError):
        api.Nlin_incompress.Compute_nlin(U)        # FFF-only entry point
    with pytest.raises(TarangGpuError):
        api.FFF_Basis.Forward_transform(b, b.real_array(), b.complex_array())  # sincos_code 0 on SFF
    f = api.FFF_Basis((16, 16, 16))
    Uf, Sf = api.FluidVF(f), api.FluidSF(f)
    with pytest.raises(TarangGpuError):
        api.Time_advance_incompress_RBC(Uf, Sf, 1e4, 1.0, 1e-3).Time_advance_step(1)
