"""The C++ host drivers (include/tarang_amd.hpp classes over the C ABI) run on
the GPU as child processes and are checked against the FP64 CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import tg_oracle as to
from gpu_util import log, rel_l2, rel_l2_vec

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "tarang_amd")
RBC_KF = (np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0))


def _run(args):
    exe = os.path.join(PKG, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (tarang_amd.build.build())")
    r = subprocess.run([exe] + [str(a) for a in args[1:]], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, f"{args[0]} exited with {r.returncode}: {r.stderr}"
    return [line.split() for line in r.stdout.splitlines() if line and not line.startswith("#")]


def test_ifluid_driver_energy_and_dissipation():
    """tarang_ifluid N steps dt nu kind: FFF NS with the seeded field, strong
    reality and corner dealias; energy and dissipation after every step."""
    N, steps, dt, nu = 32, 3, 1e-3, 0.01
    rows = _run(["tarang_ifluid", N, steps, dt, nu, 0])
    assert len(rows) == steps + 1
    o = to.Oracle((N, N, N))
    V = o.init_field(kind=0, k0=4.0, kmin=2.0, kmax=N / 3.0, energy=0.5, seed=1)
    fl = to.Fluid(nu, dt, 0, 0, 0, 0)
    for s in range(steps + 1):
        if s:
            o.rk4_step(fl, V)
        E = sum(o.energy(v) for v in V)
        D = nu * 2.0 * sum(o.Sn(v, 2) for v in V)
        e_drv, d_drv = float(rows[s][1]), float(rows[s][2])
        log(f"ifluid step {s}: E {e_drv:.15e} vs {E:.15e}, diss {d_drv:.15e} vs {D:.15e}")
        assert abs(e_drv - E) <= 1e-12 * E and abs(d_drv - D) <= 1e-12 * D


def test_rbc_driver_state_and_nusselt(tmp_path):
    """tarang_rbc Nx Ny Nz steps dt Ra Pr seed prefix: SFF RBC from u = 0 and the
    driver's theta noise; final state and Nusselt number against the oracle
    started from the driver's written initial theta."""
    N, steps, dt, Ra, Pr = (16, 16, 16), 3, 1e-3, 1e4, 1.0
    prefix = str(tmp_path / "rbc")
    rows = _run(["tarang_rbc", N[0], N[1], N[2], steps, dt, Ra, Pr, 7, prefix])
    assert len(rows) == steps
    o = to.Oracle(N, kfactor=RBC_KF, dealias=to.DEALIAS_REF_CORNER, strong_reality=False)
    cshape = o.cshape
    T0 = np.fromfile(prefix + "_T0.bin", dtype=np.complex128).reshape(cshape)
    assert np.abs(T0).max() > 0
    V = [o.czeros() for _ in range(3)]
    T = T0.copy()
    prm = to.Rbc(Ra, Pr, dt, 2)
    for s in range(steps):
        o.rk4_step_rbc(prm, V, T)
        nu_ref = o.nusselt(prm, V[0], T)
        nu_drv = float(rows[s][2])
        log(f"rbc driver step {s + 1}: Nu {nu_drv:.15e} vs {nu_ref:.15e}")
        assert abs(nu_drv - nu_ref) <= 1e-12 * abs(nu_ref)
    Vd = np.fromfile(prefix + "_V.bin", dtype=np.complex128).reshape((3,) + cshape)
    Td = np.fromfile(prefix + "_T.bin", dtype=np.complex128).reshape(cshape)
    e = rel_l2_vec([Vd[c] for c in range(3)] + [Td], V + [T])
    log(f"rbc driver final state rel L2 {e:.3e} (|u| {np.abs(Vd).max():.3e})")
    assert np.abs(Vd).max() > 0 and e < 1e-12
    assert rel_l2(Td, T) < 1e-12
