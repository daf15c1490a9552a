"""The C ABI library loads, exports every entry point include/psgpu.h declares,
and the product package does not depend on the oracle (CPU only, no GPU calls)."""
import os
import re
import subprocess

import pytest

from puresoft3d_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    names = N.declared_symbols()
    assert len(names) >= 50
    L = N.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (psgpu_\w+)", out))
    assert set(names) <= exported
    # the ctypes signatures cover the whole header
    assert set(names) == set(N._SIGS)


def test_library_is_gfx950_code_object():
    """the offload bundle of the library built here names the gfx950 target"""
    data = open(N.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_serving_library_exports_psserve_header():
    hdr = open(os.path.join(ROOT, "include", "psserve.h")).read()
    names = set(re.findall(r"^\s*(?:int|const char\*)\s+(psserve_\w+)\s*\(", hdr, re.M))
    assert names == {"psserve_create", "psserve_create_multi", "psserve_set_config", "psserve_destroy", "psserve_run",
                     "psserve_last_error"}
    lib = os.path.join(ROOT, "puresoft3d_amd", "libpuresoft3d.so")
    if not os.path.exists(lib):
        pytest.skip("host library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    assert names <= set(re.findall(r" T (psserve_\w+)", out))


def test_product_code_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "puresoft3d_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dirpath, fn)).read()
                assert "pyoracle" not in txt and "ps_oracle" not in txt and "libpsoracle" not in txt, fn


def test_bench_uses_oracle_only_in_cpu_baseline():
    src = open(os.path.join(ROOT, "bench.py")).read()
    head, _, rest = src.partition("def cpu_baseline(")
    assert "pyoracle" not in head
    body = rest.split("\ndef ", 1)[1]
    assert "pyoracle" not in body


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(N, "_lib", None)
    with pytest.raises(ImportError):
        N.lib()
