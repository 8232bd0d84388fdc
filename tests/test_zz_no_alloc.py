"""tg_shell_spectrum_noalloc (the allocation-free variant of
tg_shell_spectrum, static device scratch) against the plan-based entry point:
bit-identical spectra on two calls in a row, the second on a smaller grid;
grids beyond the static scratch are rejected before any device work."""
import ctypes

import numpy as np
import pytest

import tg_oracle as to
from gpu_util import dev, log

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from tarang_amd import api  # noqa: E402
from tarang_amd._lib import ERR_ARG, TgConfig  # noqa: E402


def _cfg(N, kf=(1.0, 1.0, 1.0)):
    return TgConfig(0, N[0], N[1], N[2], (ctypes.c_double * 3)(*kf), 0, 1, 1, 0)


def _spectra(N, seed, n, kf=(1.0, 1.0, 1.0)):
    o = to.Oracle(N, kfactor=kf)
    V = o.init_field(kind=1, kmin=1.0, kmax=N[0] / 3, energy=0.5, seed=seed)
    b = api.FFF_Basis(N, kfactor=kf)
    A = dev(V[1])
    ref = b.Compute_shell_spectrum(A, n)
    nb = len(ref) + 3
    out = torch.full((nb,), -1.0, dtype=torch.float64, device="cuda")
    cfg = _cfg(N, kf)
    s = torch.cuda.current_stream()
    rc = b.L.tg_shell_spectrum_noalloc(ctypes.byref(cfg), ctypes.c_void_p(A.data_ptr()), n,
                                       ctypes.c_void_p(out.data_ptr()), nb, ctypes.c_void_p(s.cuda_stream))
    assert rc == 0, b.L.tg_last_error_string()
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    return got, ref


def test_noalloc_spectrum_matches_plan_entry_point():
    got1, ref1 = _spectra((64, 64, 64), 81, 0)
    log(f"noalloc 64^3: {len(ref1)} bins, max |diff| {np.abs(got1[:len(ref1)] - ref1).max():.1e}")
    assert np.array_equal(got1[:len(ref1)], ref1) and np.all(got1[len(ref1):] == 0)
    # second call, smaller input: the shared scratch keeps the first call's
    # contents, which must not leak into this result
    got2, ref2 = _spectra((32, 16, 32), 82, 2, kf=(1.0, 2.0, 1.0))
    log(f"noalloc 32x16x32: {len(ref2)} bins, max |diff| {np.abs(got2[:len(ref2)] - ref2).max():.1e}")
    assert np.array_equal(got2[:len(ref2)], ref2) and np.all(got2[len(ref2):] == 0)


def test_noalloc_rejects_oversized_grids():
    b = api.FFF_Basis((16, 16, 16))
    A = b.complex_array()
    out = torch.zeros(2000, dtype=torch.float64, device="cuda")
    cfg = _cfg((1024, 1024, 1024))
    rc = b.L.tg_shell_spectrum_noalloc(ctypes.byref(cfg), ctypes.c_void_p(A.data_ptr()), 0,
                                       ctypes.c_void_p(out.data_ptr()), 2000, None)
    assert rc == ERR_ARG and b"scratch" in b.L.tg_last_error_string()
    cfg = _cfg((16, 16, 16))
    rc = b.L.tg_shell_spectrum_noalloc(ctypes.byref(cfg), ctypes.c_void_p(A.data_ptr()), 0,
                                       ctypes.c_void_p(out.data_ptr()), 5, None)
    assert rc == ERR_ARG
