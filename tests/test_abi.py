"""CPU-side checks of the C ABI: the library loads, exports every function
that include/ugemm_hip.h declares, and rejects bad arguments before it touches
a device (no compute calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import ugemm_amd as ug

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ugemm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*\b([a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("sgemm_hip", "dgemm_hip", "ugemm_hip_init", "ugemm_hip_finish", "sgemm_hip_packed",
                 "ugemm_hip_sgemm_dev", "ugemm_hip_dgemm_dev", "ugemm_hip_sgemm_grouped",
                 "ugemm_hip_last_error", "ugemm_hip_tune_load", "ugemm_hip_tune_save"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(ug.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    ug.load()   # argtypes for everything the bindings use


def test_status_strings_and_version():
    L = ug.load()
    assert L.ugemm_hip_status_string(0) == b"ok"
    assert L.ugemm_hip_status_string(ug.ERR_ARG) == b"invalid argument"
    assert b"gfx950" in L.ugemm_hip_version()


@pytest.mark.parametrize("args,msg", [
    (("X", "N", "N", 4, 4, 4, 4, 4, 4), "major"),
    (("R", "Q", "N", 4, 4, 4, 4, 4, 4), "transa"),
    (("R", "N", "N", -1, 4, 4, 4, 4, 4), "negative"),
    (("R", "N", "N", 4, 4, 8, 7, 4, 4), "lda"),      # row-major N: lda >= K
    (("R", "T", "N", 9, 4, 4, 8, 4, 4), "lda"),      # row-major T: lda >= M
    (("R", "N", "T", 4, 4, 6, 6, 5, 4), "ldb"),      # row-major T B: ldb >= K
    (("R", "N", "N", 4, 6, 4, 4, 6, 5), "ldc"),      # ldc >= N
    (("C", "N", "N", 6, 4, 4, 5, 4, 6), "lda"),      # column-major N: lda >= M
    (("C", "N", "N", 4, 6, 4, 4, 4, 3), "ldc"),      # column-major: ldc >= M
])
def test_argument_validation_without_device(args, msg):
    major, ta, tb, M, N, K, lda, ldb, ldc = args
    buf = np.zeros(1024, np.float32)
    with pytest.raises(ug.UgemmError) as e:
        ug.sgemm(major, ta, tb, M, N, K, 1.0, buf, lda, buf, ldb, 0.0, buf, ldc)
    assert e.value.status == ug.ERR_ARG
    assert msg in str(e.value)


def test_unknown_mode_rejected():
    buf = np.zeros(64, np.float32)
    with pytest.raises(ug.UgemmError) as e:
        ug.sgemm("R", "N", "N", 4, 4, 4, 1.0, buf, 4, buf, 4, 0.0, buf, 4, mode=17)
    assert e.value.status == ug.ERR_ARG


def test_null_operands_rejected():
    buf = np.zeros(64, np.float32)
    with pytest.raises(ug.UgemmError) as e:
        ug.sgemm("R", "N", "N", 4, 4, 4, 1.0, None, 4, buf, 4, 0.0, buf, 4)
    assert e.value.status == ug.ERR_ARG
    with pytest.raises(ug.UgemmError):
        ug.sgemm("R", "N", "N", 4, 4, 4, 1.0, buf, 4, buf, 4, 0.0, None, 4)


def test_void_entry_records_status():
    L = ug.load()
    buf = np.zeros(64, np.float32)
    L.sgemm_hip(b"R", b"N", b"N", 4, 4, 8, 1.0, buf.ctypes.data, 7, buf.ctypes.data, 4, 0.0, buf.ctypes.data, 4)
    assert L.ugemm_hip_last_error() == ug.ERR_ARG
    assert b"lda" in L.ugemm_hip_last_error_message()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 (or a multi-rank launch) exits
    with status 2 and says why, before touching any device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2 and "one GPU" in r.stderr
    env = dict(os.environ, WORLD_SIZE="2")
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py")], capture_output=True, text=True,
                       timeout=120, cwd=root, env=env)
    assert r.returncode == 2 and "one GPU" in r.stderr
