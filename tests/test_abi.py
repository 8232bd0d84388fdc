"""The C ABI library loads without a GPU, exports every entry point declared
in include/tarang_gpu.h, and rejects bad arguments before touching a device.
CPU only: no compute call is made."""
import ctypes
import os
import subprocess

import pytest

from tarang_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    L = _lib.load()
    syms = _lib.header_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    # every exported tg_* symbol is declared in the header
    extra = sorted(s for s in exported if s.startswith("tg_") and s not in syms)
    assert not extra, extra


def test_binding_covers_header():
    assert set(_lib.header_symbols()) <= set(_lib.SIGNATURES)


def _cfg(**kw):
    c = dict(basis=0, nx=32, ny=32, nz=32, kfactor=(1.0, 1.0, 1.0), precision=0, dealias=1, strong_reality=1, device=0)
    c.update(kw)
    return _lib.TgConfig(c["basis"], c["nx"], c["ny"], c["nz"], (ctypes.c_double * 3)(*c["kfactor"]), c["precision"],
                         c["dealias"], c["strong_reality"], c["device"])


@pytest.mark.parametrize("kw,code", [
    (dict(nx=48, precision=1), -2), (dict(nx=56), -2), (dict(ny=72), -2), (dict(nx=12), -2), (dict(nz=8), -2),
    (dict(nz=36), -2), (dict(nz=1280), -2), (dict(ny=2048), -2), (dict(basis=3), -1), (dict(precision=7), -1),
    (dict(dealias=5), -1), (dict(kfactor=(1.0, 0.0, 1.0)), -1),
])
def test_plan_create_rejects_bad_config(kw, code):
    L = _lib.load()
    cfg = _cfg(**kw)
    h = ctypes.c_void_p()
    rc = L.tg_plan_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert L.tg_last_error_string()


def test_null_arguments_are_errors():
    L = _lib.load()
    assert L.tg_plan_create(None, None) == -1
    assert L.tg_forward(None, 0, None, None) == -1
    assert L.tg_rk4_step(None, None, None, 1, 0) == -1
    assert L.tg_plan_destroy(None) == 0
    assert L.tg_rhs_vjp(None, None, None, None) == -1
    assert L.tg_adjoint_forward(None, None, None, 1, 1) == -1
    assert L.tg_adjoint_backward(None, None) == -1
    assert L.tg_adjoint_free(None) == -1
    assert L.tg_kernel_name(3) == b"x_nlin_rk"
    assert L.tg_kernel_name(11) == b"z_adj"
    assert L.tg_version().startswith(b"tarang_amd")


def test_header_is_c_compatible(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "tarang_gpu.h"\nint main(void){tg_config c; (void)c; return 0;}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(tmp_path / "t")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_io_library_exports_every_header_symbol():
    import re
    from tarang_amd import h5io
    h5io.load()
    text = open(os.path.join(ROOT, "include", "tarang_io.h")).read()
    syms = sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(tgio_\w+)\s*\(", text, flags=re.M)))
    assert len(syms) == 4
    out = subprocess.run(["nm", "-D", "--defined-only", h5io.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(syms) <= exported
    assert not sorted(s for s in exported if s.startswith("tgio_") and s not in syms)
