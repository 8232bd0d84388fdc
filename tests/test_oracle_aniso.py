"""Pins of the oracle's anisotropic diagnostics (oracle/tg_oracle_aniso.c):
polar angles against numpy's geometry, single modes in their bins, sector
sums against shell sums, ring-to-ring transfers summed over sectors against
the (separately pinned) shell-to-shell transfer, antisymmetry under the 2/3
rule.  CPU only."""
import numpy as np
import pytest

import tg_oracle as to


def kvec(o):
    kx = np.fft.fftfreq(o.nx, 1.0 / o.nx)
    ky = np.fft.fftfreq(o.ny, 1.0 / o.ny)
    kx[o.nx // 2] = o.nx // 2
    ky[o.ny // 2] = o.ny // 2
    kz = np.arange(o.nz // 2 + 1)
    KX, KY, KZ = np.meshgrid(kx * o.kfactor[0], ky * o.kfactor[1], kz * o.kfactor[2], indexing="ij")
    return KX, KY, KZ


def mult(o):
    return np.array([[[o.multiplicity(lx, ly, lz) for lz in range(o.nz // 2 + 1)] for ly in range(o.ny)]
                     for lx in range(o.nx)])


def equi_angles(nsec):
    """EQUISPACED sectors over [0, pi/2] (Global_process_advanced_vars.cc:100-107)."""
    a = np.array([i * (np.pi / 2) / nsec for i in range(nsec)] + [np.pi / 2])
    return a


@pytest.mark.parametrize("dirn", [1, 2, 3])
def test_polar_angle_geometry(dirn):
    o = to.Oracle((8, 6, 10), kfactor=(1.0, 0.5, 2.0))
    K = kvec(o)
    Kmag = np.sqrt(K[0] ** 2 + K[1] ** 2 + K[2] ** 2)
    kpll = K[dirn - 1]
    theta = np.where(Kmag > 0, np.arccos(np.clip(kpll / np.where(Kmag > 0, Kmag, 1), -1, 1)), 0.0)
    for lx in range(8):
        for ly in range(6):
            for lz in range(6):
                assert abs(o.polar_angle(lx, ly, lz, dirn) - theta[lx, ly, lz]) < 1e-14


def test_ring_spectrum_single_mode_and_sector_sums():
    o = to.Oracle(16)
    nsec = 5
    ang = equi_angles(nsec)
    A = [o.czeros() for _ in range(3)]
    A[0][2, 3, 1] = 0.5 + 0.25j          # K = (2, 3, 1), |K| = 3.74 -> shell 4
    R = o.ring_spectrum(A, 0, ang, dirn=3)
    th = np.arccos(1 / np.sqrt(14.0))     # to the z axis
    sec = next(i for i in range(nsec + 1) if ang[i] >= th)
    nz = np.argwhere(R[0] != 0)
    assert nz.tolist() == [[4, sec]]
    assert abs(R[0, 4, sec] - 0.3125) < 1e-16 and not R[1].any() and not R[2].any()
    # random field: the sector sum is the shell sum m K^n |A|^2 over 0 < K, ceil(K) <= Max_radius_inside
    rng = np.random.default_rng(1)
    B = [(rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)) for _ in range(3)]
    KX, KY, KZ = kvec(o)
    Km = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    m = mult(o)
    kmax = o.max_radius_inside()
    sh = np.ceil(Km).astype(int)
    ok = (Km > 0) & (sh <= kmax)
    for n in (0, 2):
        R = o.ring_spectrum(B, n, ang, dirn=2)
        for c in range(3):
            ref = np.bincount(sh[ok], weights=(m * Km ** n * np.abs(B[c]) ** 2)[ok], minlength=kmax + 1)
            assert np.abs(R[c].sum(axis=1) - ref).max() <= 1e-12 * ref.max()
            assert not R[c][:, 0].any()      # sector 0 holds nothing (theta > 0 or sector 1)


def test_cyl_ring_spectrum_single_modes():
    o = to.Oracle(16)
    kp = np.linspace(0.0, 8.0, 5)           # slabs (0,2], (2,4], ... of |Kpll|
    V = [o.czeros() for _ in range(3)]
    V[2][3, 0, 3] = 1.0                     # along z (dirn 3): Kperp = 3, Kpll = 3 -> slab 2
    V[0][0, 5, 0] = 2.0                     # perpendicular, Kpll = 0 -> slab 1, Kperp = 5
    S1, S2 = o.cyl_ring_spectrum(V, 0, kp, dirn=3)
    assert np.argwhere(S1 != 0).tolist() == [[3, 2]] and S1[3, 2] == 1.0
    # the kz = 0 mode has multiplicity 1/2
    assert np.argwhere(S2 != 0).tolist() == [[5, 1]] and S2[5, 1] == 2.0
    assert o.aniso_max_krho_inside(3) == 8


def test_fill_ring_sectors_tile_the_shell():
    o = to.Oracle(12)
    rng = np.random.default_rng(2)
    A = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    ang = equi_angles(4)
    acc = o.czeros()
    for j in range(1, 5):
        acc += o.fill_array_ring(A, 2.0, 5.0, ang[j - 1], ang[j], dirn=3)
    assert np.array_equal(acc, o.fill_array_shell(A, 2.0, 5.0))


def test_ring_to_ring_sums_to_shell_to_shell():
    """dirn 3: every stored mode has theta in [0, pi/2], so the sector rings
    of a shell tile it and the transfer summed over giver and receiver
    sectors is the shell-to-shell transfer."""
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    V = o.init_field(kind=1, kmin=1.0, kmax=6.0, energy=0.5, seed=3)
    radii = np.array([0.0, 2.0, 4.0, 6.0, np.inf])
    ang = equi_angles(3)
    T = o.ring_to_ring([v.copy() for v in V], radii, ang, dirn=3)
    S = o.shell_to_shell([v.copy() for v in V], radii)
    TS = T.sum(axis=(1, 3))
    assert np.abs(TS - S).max() <= 1e-13 * np.abs(S).max()
    assert np.abs(S).max() > 1e-4


def test_ring_to_ring_antisymmetric_two_thirds():
    o = to.Oracle(16, dealias=to.DEALIAS_TWOTHIRDS)
    V = o.init_field(kind=1, kmin=1.0, kmax=5.0, energy=0.5, seed=4)
    for v in V:
        o.dealias(v)
    radii = np.array([0.0, 2.0, 4.0, 6.0, np.inf])
    ang = equi_angles(3)
    T = o.ring_to_ring([v.copy() for v in V], radii, ang, dirn=3)
    ns, na = len(radii) - 1, len(ang) - 1
    G = T[1:ns, 1:, 1:ns, 1:].reshape((ns - 1) * na, (ns - 1) * na)
    assert np.abs(G + G.T).max() <= 1e-13 * np.abs(G).max()
    assert abs(T.sum()) <= 1e-13 * np.abs(T).sum()


def test_cyl_ring_to_ring_slab_sums():
    """Summed over giver slabs (they tile Kpll >= 0 for dirn 3), the giver is
    the |K| shell; summed over receiver slabs, the receivers are the Kperp
    shells: compare with the shell-filled giver term binned by Kperp."""
    o = to.Oracle(16, dealias=to.DEALIAS_REF_CORNER)
    V = o.init_field(kind=1, kmin=1.0, kmax=6.0, energy=0.5, seed=5)
    radii = np.array([0.0, 2.0, 4.0, 6.0, np.inf])
    kp = np.linspace(0.0, 8.0, 4)
    T = o.cyl_ring_to_ring([v.copy() for v in V], radii, kp, dirn=3)
    Vd = [v.copy() for v in V]
    for v in Vd:
        o.dealias(v)
    KX, KY, KZ = kvec(o)
    Kperp = np.sqrt(KX ** 2 + KY ** 2)
    m = mult(o)
    kmax = o.aniso_max_krho_inside(3)
    ok = Kperp <= kmax
    shell = np.searchsorted(radii, Kperp, side="left")
    for i in range(1, len(radii) - 1):
        W = [o.fill_array_shell(v, radii[i - 1], radii[i]) for v in Vd]
        N = o.compute_nlin_giver([v.copy() for v in Vd], W)
        d = sum((N[c] * np.conj(Vd[c])).real for c in range(3))
        ref = -np.bincount(shell[ok], weights=(2 * m * d)[ok], minlength=len(radii))
        got = T[i].sum(axis=(0, 2))
        assert np.abs(got - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1e-30), (i, got, ref)
