"""The C-ABI library builds, loads and exports every entry point declared in
include/caviar_amd.h; without a GPU it refuses politely (no compute calls)."""
import ctypes
import os
import subprocess

import pytest

import caviar_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib_path():
    if not os.path.exists(caviar_amd.LIB_PATH):
        from caviar_amd import build
        build.build()
    return caviar_amd.LIB_PATH


def test_library_exports_every_declared_symbol(lib_path):
    lib = ctypes.CDLL(lib_path)
    declared = caviar_amd.exported_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= exported


def test_library_is_gfx950_only(lib_path):
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", lib_path], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout
    assert "gfx90a" not in out.stdout and "gfx942" not in out.stdout


def test_engine_requires_the_library_not_a_fallback(monkeypatch, tmp_path):
    monkeypatch.setattr(caviar_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(caviar_amd, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        caviar_amd.load_library()


def test_create_without_device_returns_an_error_code(lib_path):
    try:
        import torch  # noqa: F401
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    lib = caviar_amd.load_library()
    h = ctypes.c_void_p()
    rc = lib.cav_create(0, ctypes.byref(h))
    assert rc != 0 and not h.value
    # null handles are rejected, not dereferenced
    assert lib.cav_setup(None) != 0
    assert lib.cav_step(None, 0.001, 1) != 0
