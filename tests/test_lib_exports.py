"""The C-ABI library loads and exports every function include/*.h declares (no GPU needed)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pfem_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_functions():
    names = declared_functions()
    assert "pfem_pspg_step" in names and "pfem_wc_step" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from pfem_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libpfem_gpu.so is not built (run python -m pfem_amd.build)")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"missing exports: {missing}"
    assert set(_lib.EXPORTS) <= set(declared_functions())


def test_loader_fails_loudly_without_library(tmp_path, monkeypatch):
    from pfem_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no fallback"):
        _lib.lib()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asking for more exits with status 2 and says why, before
    touching a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stdout + r.stderr
