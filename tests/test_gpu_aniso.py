"""Anisotropic diagnostics on the GPU against the FP64 oracle (whose pins are
in test_oracle_aniso.py): ring and cylindrical-ring spectra, ring-to-ring
and cylindrical-ring-to-ring transfers, for each anisotropy axis."""
import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, log, rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402

KF = (1.0, 0.5, 2.0)
N = (32, 16, 32)


def field(o, seed):
    return o.init_field(kind=1, kmin=1.0, kmax=10.0, energy=0.5, seed=seed)


@pytest.mark.parametrize("dirn", [1, 2, 3])
def test_ring_spectra(dirn):
    o = to.Oracle(N, kfactor=KF)
    V = field(o, 71)
    b = api.FFF_Basis(N, kfactor=KF)
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    ang = api.sector_angles_equi(5)
    for n in (0, 2):
        got = api.Correlation.Compute_ring_spectrum(U, n, ang, dirn)
        ref = o.ring_spectrum([v.copy() for v in V], n, ang, dirn)
        e = rel_l2(got, ref)
        log(f"ring spectrum dirn={dirn} n={n}: rel L2 {e:.2e}")
        assert got.shape == ref.shape and e < 1e-13
    T = api.FluidSF(b)
    T.set_field(dev(V[0]))
    got = api.Correlation.Compute_ring_spectrum(T, 0, ang, dirn)
    assert rel_l2(got[0], o.ring_spectrum([V[0].copy()], 0, ang, dirn)[0]) < 1e-13
    kp = api.kpll_edges_equi(b, 4, dirn)
    S1, S2 = api.Correlation.Compute_cylindrical_ring_spectrum(U, 0, kp, dirn)
    R1, R2 = o.cyl_ring_spectrum([v.copy() for v in V], 0, kp, dirn)
    log(f"cylindrical spectrum dirn={dirn}: {rel_l2(S1, R1):.2e} {rel_l2(S2, R2):.2e}")
    assert rel_l2(S1, R1) < 1e-13 and rel_l2(S2, R2) < 1e-13


@pytest.mark.parametrize("dirn", [1, 3])
def test_ring_transfers(dirn):
    Ns = (16, 16, 16)
    o = to.Oracle(Ns, dealias=to.DEALIAS_REF_CORNER)
    V = o.init_field(kind=1, kmin=1.0, kmax=6.0, energy=0.5, seed=72)
    b = api.FFF_Basis(Ns)
    U = api.FluidVF(b)
    U.set_field([dev(v) for v in V])
    radii = np.array([0.0, 2.0, 4.0, 6.0, np.inf])
    ang = api.sector_angles_equi(3)
    et = api.EnergyTr(b)
    got = et.Compute_ring_tr(U, radii, ang, dirn)
    ref = o.ring_to_ring([v.copy() for v in V], radii, ang, dirn)
    e = rel_l2(got, ref)
    kp = api.kpll_edges_equi(b, 3, dirn)
    gotc = et.Compute_cylindrical_ring_tr(U, radii, kp, dirn)
    refc = o.cyl_ring_to_ring([v.copy() for v in V], radii, kp, dirn)
    ec = rel_l2(gotc, refc)
    log(f"ring-to-ring dirn={dirn}: rel L2 {e:.2e}; cylindrical {ec:.2e}")
    assert e < 1e-11 and ec < 1e-11
    assert np.abs(ref).max() > 1e-5 and np.abs(refc).max() > 1e-5
