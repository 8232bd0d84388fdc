"""Field dumps in the reference's HDF5 layout (libtarang_io.so), on host
arrays.  Pinned by an independent reader (the HDF5 distribution's h5dump),
by a numpy restatement of the reduced-plan filters of
BasicIO::Set_H5_plans (BasicIO_SP.h:159-214), by round trips, and by the
incompressibility rebuild of V3 (Universal::Fill_Vz)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from tarang_amd import h5io

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def rand_c(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex128)


def test_full_round_trip_and_names(tmp_path):
    N = (8, 6, 10)
    A = rand_c((8, 6, 6), 1)
    h5io.write(A, N, h5io.FULL, str(tmp_path), "time_0.5", "U.V1")
    path = tmp_path / "time_0.5" / "U.V1.h5"
    assert path.exists()
    B = np.zeros_like(A)
    h5io.read(B, N, h5io.FULL, str(tmp_path / "time_0.5"), "U.V1")
    assert np.array_equal(A, B)
    assert h5io.time_folder(0.5) == "time_0.5" and h5io.time_folder(1e-5) == "time_1e-05"
    assert h5io.time_folder(2.0) == "time_2" and h5io.time_folder(1234567.0) == "time_1.23457e+06"


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")
def test_h5dump_reads_the_dataset(tmp_path):
    """h5dump (not this library) sees one dataset named like the file, of
    shape (Nx, Ny, Nz/2+1), a compound of two IEEE doubles r and i, with
    the array's values in row-major order."""
    N = (2, 2, 4)
    A = (np.arange(2 * 2 * 3) + 1j * (100 + np.arange(2 * 2 * 3))).reshape(2, 2, 3) / 8.0
    h5io.write(A, N, h5io.FULL, str(tmp_path), "time_0", "T.F")
    out = subprocess.run([H5DUMP, str(tmp_path / "time_0" / "T.F.h5")], capture_output=True, text=True,
                         timeout=60).stdout
    assert 'DATASET "T.F"' in out
    assert "DATASPACE  SIMPLE { ( 2, 2, 3 ) / ( 2, 2, 3 ) }" in out
    assert re.search(r'H5T_IEEE_F64LE "r";\s*H5T_IEEE_F64LE "i";', out)
    # h5dump prints "(i,j,k): {re, im}" entries in row-major order
    body = out.split("DATA {", 1)[1]
    pairs = re.findall(r"\{\s*([-+0-9.e]+),\s*([-+0-9.e]+)\s*\}", body)
    got = np.array([float(a) + 1j * float(b) for a, b in pairs])
    assert got.size == A.size and np.array_equal(got, A.ravel())


def reduced_ref(A, N, R, fourier_x):
    """The reduced file array from the filters of BasicIO_SP.h:183-196: a
    Fourier axis keeps file [0, R/2) <- memory [0, R/2) and file [R/2+1, R)
    <- memory [N-R/2+1, N) (file R/2 left zero); x of SFF and z keep the
    first entries."""
    nx, ny = N[0], N[1]
    rz = R[2] // 2 + 1
    out = np.zeros((R[0], R[1], rz), dtype=A.dtype)

    def sel(n, r, fourier):
        if not fourier:
            return list(range(r)), list(range(r))
        f = list(range(r // 2)) + list(range(r // 2 + 1, r))
        m = list(range(r // 2)) + list(range(n - r // 2 + 1, n))
        return f, m

    fx, mx = sel(nx, R[0], fourier_x)
    fy, my = sel(ny, R[1], True)
    out[np.ix_(fx, fy, range(rz))] = A[np.ix_(mx, my, range(rz))]
    return out


@pytest.mark.parametrize("fourier_x", [True, False])
def test_reduced_write_and_read(tmp_path, fourier_x):
    N, R = (16, 8, 12), (8, 4, 6)
    A = rand_c((16, 8, 7), 2)
    h5io.write(A, N, h5io.REDUCED, str(tmp_path), "reduced_1", "U.V2", fourier_x, R)
    # read the file back as a full array of the reduced grid: it must equal the restatement
    F = np.zeros((8, 4, 4), dtype=np.complex128)
    h5io.read(F, R, h5io.FULL, str(tmp_path / "reduced_1"), "U.V2")
    assert np.array_equal(F, reduced_ref(A, N, R, fourier_x))
    # the reduced read zeroes the array and fills the kept modes
    B = rand_c((16, 8, 7), 3)
    h5io.read(B, N, h5io.REDUCED, str(tmp_path / "reduced_1"), "U.V2", fourier_x, R)
    mask = np.zeros(A.shape, dtype=bool)
    nx, ny = N[0], N[1]
    xs = list(range(R[0] // 2)) + list(range(nx - R[0] // 2 + 1, nx)) if fourier_x else list(range(R[0]))
    ys = list(range(R[1] // 2)) + list(range(ny - R[1] // 2 + 1, ny))
    mask[np.ix_(xs, ys, range(R[2] // 2 + 1))] = True
    assert np.array_equal(B[mask], A[mask]) and np.all(B[~mask] == 0)


def test_kz0_plane_and_fill_vz(tmp_path):
    """V1, V2 and V3's kz = 0 plane written (output_vx_vy_switch); read back
    with V3 rebuilt from incompressibility: the original V3 is recovered."""
    N = (8, 8, 8)
    kf = (1.0, 2.0, 0.5)
    kx = np.fft.fftfreq(8, 1 / 8)
    kx[4] = 4
    KX, KY, KZ = np.meshgrid(kx * kf[0], kx * kf[1], np.arange(5) * kf[2], indexing="ij")
    V1, V2 = rand_c((8, 8, 5), 4), rand_c((8, 8, 5), 5)
    V3 = rand_c((8, 8, 5), 6)
    V3[:, :, 1:] = -(KX * V1 + KY * V2)[:, :, 1:] / KZ[:, :, 1:]
    for name, a, plan in (("U.V1", V1, h5io.FULL), ("U.V2", V2, h5io.FULL), ("U.V3kz0", V3, h5io.KZ0_FULL)):
        h5io.write(a, N, plan, str(tmp_path), "time_1", name)
    W3 = np.zeros_like(V3)
    h5io.read(W3, N, h5io.KZ0_FULL, str(tmp_path / "time_1"), "U.V3kz0")
    assert np.array_equal(W3[:, :, 0], V3[:, :, 0]) and np.all(W3[:, :, 1:] == 0)
    h5io.fill_vz(V1, V2, W3, N, kf)
    assert np.abs(W3 - V3).max() <= 1e-15 * np.abs(V3).max() * 10
    # the rebuilt field is divergence-free for kz > 0
    div = (KX * V1 + KY * V2 + KZ * W3)[:, :, 1:]
    assert np.abs(div).max() < 1e-13


def test_real_plan_and_errors(tmp_path):
    N = (4, 4, 6)
    r = np.random.default_rng(7).standard_normal((4, 4, 8))
    h5io.write(r, N, h5io.REAL, str(tmp_path), "time_0", "U.V1r")
    s = np.zeros_like(r)
    h5io.read(s, N, h5io.REAL, str(tmp_path / "time_0"), "U.V1r")
    assert np.array_equal(r, s)
    # dimensions differing from the plan: the reference's message, TG_ERR_SHAPE
    A = np.zeros((8, 4, 4), dtype=np.complex128)
    with pytest.raises(h5io.TarangIoError) as ei:
        h5io.read(A, (8, 4, 6), h5io.FULL, str(tmp_path / "time_0"), "U.V1r")
    assert ei.value.code == -2 and "differ from specified dimensions" in str(ei.value)
    with pytest.raises(h5io.TarangIoError) as ei:
        h5io.read(A, (8, 4, 6), h5io.FULL, str(tmp_path), "missing")
    assert ei.value.code == -1 and "File does not exist" in str(ei.value)
    with pytest.raises(h5io.TarangIoError):
        h5io.write(A, (8, 4, 6), h5io.REDUCED, str(tmp_path), "x", "y", True, (16, 4, 6))
