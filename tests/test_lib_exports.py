"""The C-ABI library loads and exports every function include/acumen.h declares
(no GPU calls)."""
import ctypes
import os
import re

import pytest

from acumen_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(acu_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("acu_plan_create", "acu_radon", "acu_iradon", "acu_get_slice", "acu_tvopt",
                 "acu_batch_tvopt", "acu_batch_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libacumen_hip.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_library_reports_missing_device_cleanly():
    """Without a GPU, creating a context fails with a status, not a crash."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libacumen_hip.so not built")
    import subprocess
    import sys
    code = ("import ctypes; from acumen_amd import _lib; L=_lib.load(); h=ctypes.c_void_p();"
            "rc=L.acu_ctx_create(0, ctypes.byref(h)); print(rc, L.acu_last_error().decode())")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT,
                         env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    rc = int(out.stdout.split()[0])
    assert rc == _lib.ACU_ERR_HIP


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(FileNotFoundError):
        _lib.load(str(tmp_path / "libacumen_hip.so"))
