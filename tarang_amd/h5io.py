"""Field dumps in the reference's HDF5 layout (ctypes over libtarang_io.so,
include/tarang_io.h).

Arrays are numpy arrays (host memory) or torch tensors on the GPU (device
memory); complex arrays have the spectral shape (Nx, Ny, Nz/2+1), real arrays
(Nx, Ny, Nz+2).  File and dataset naming follow BasicIO::Write / Read
(trunk/lib/basis/basicfn/BasicIO.cc:222-312): <dir>/<folder>/<name>.h5 with a
dataset <name>.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtarang_io.so")

FULL, KZ0_FULL, REDUCED, KZ0_REDUCED, REAL = 0, 1, 2, 3, 4


class TgioLayout(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("fourier_x", ctypes.c_int),
                ("reduced", ctypes.c_int * 3), ("on_device", ctypes.c_int)]


class TarangIoError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with code {code}: {msg}")
        self.code = code


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run tarang_amd.build.build() first")
        lib = ctypes.CDLL(LIB_PATH)
        D, P, I, C = ctypes.c_void_p, ctypes.POINTER(TgioLayout), ctypes.c_int, ctypes.c_char_p
        lib.tgio_write.argtypes = [D, P, I, C, C, C]
        lib.tgio_read.argtypes = [D, P, I, C, C]
        lib.tgio_fill_vz.argtypes = [D, D, D, P, ctypes.POINTER(ctypes.c_double)]
        for f in (lib.tgio_write, lib.tgio_read, lib.tgio_fill_vz):
            f.restype = I
        lib.tgio_last_error_string.restype = C
        _lib = lib
    return _lib


def _check(fn, rc):
    if rc != 0:
        raise TarangIoError(fn, rc, load().tgio_last_error_string().decode())


def _ptr(a):
    """(pointer, on_device) of a C-contiguous numpy array or torch tensor.
    Device work queued on torch's streams is waited for first (the library
    copies with the blocking hipMemcpy)."""
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data, 0
    if not a.is_contiguous():
        raise ValueError("tensor must be contiguous")
    if a.is_cuda:
        import torch
        torch.cuda.synchronize(a.device)
        return a.data_ptr(), 1
    return a.data_ptr(), 0


def layout(N, fourier_x=True, reduced=None, on_device=False):
    r = reduced if reduced is not None else (0, 0, 0)
    return TgioLayout(int(N[0]), int(N[1]), int(N[2]), 1 if fourier_x else 0, (ctypes.c_int * 3)(*r),
                      1 if on_device else 0)


def write(a, N, plan, out_dir, folder, name, fourier_x=True, reduced=None):
    """BasicIO::Write of one array with an IO plan (FULL, KZ0_FULL, REDUCED,
    KZ0_REDUCED, REAL)."""
    p, dev = _ptr(a)
    lay = layout(N, fourier_x, reduced, dev)
    _check("tgio_write", load().tgio_write(p, ctypes.byref(lay), plan, out_dir.encode(), folder.encode(),
                                           name.encode()))


def read(a, N, plan, in_dir, name, fourier_x=True, reduced=None):
    """BasicIO::Read into `a` (same plans)."""
    p, dev = _ptr(a)
    lay = layout(N, fourier_x, reduced, dev)
    _check("tgio_read", load().tgio_read(p, ctypes.byref(lay), plan, in_dir.encode(), name.encode()))
    return a


def fill_vz(V1, V2, V3, N, kfactor=(1.0, 1.0, 1.0)):
    """Universal::Fill_Vz (FFF): V3 for kz > 0 from incompressibility."""
    p1, d1 = _ptr(V1)
    p2, _ = _ptr(V2)
    p3, _ = _ptr(V3)
    lay = layout(N, True, None, d1)
    kf = (ctypes.c_double * 3)(*kfactor)
    _check("tgio_fill_vz", load().tgio_fill_vz(p1, p2, p3, ctypes.byref(lay), kf))
    return V3


def time_folder(t):
    """Folder of a dump at time t: "time_" + To_string(t) (def_vars.h:164-169,
    the stream's default 6-significant-digit format)."""
    return "time_" + format(float(t), "g")
