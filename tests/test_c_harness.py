"""The C drop-in: tests/c_harness/check_sgemm_hip.c calls sgemm_hip / dgemm_hip
through the uut function-pointer type of the reference's harness
(check_sgemm.c:87-103, check_dgemm.c:86-102) and the packed sgemm_ocl-style
entry (sgemm_ocl2.h:166), and checks every result against a long-double loop."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_harness", "check_sgemm_hip.c")
LIBDIR = os.path.join(ROOT, "ugemm_amd")


def _build(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libugemm_hip.so")):
        pytest.skip("libugemm_hip.so is not built")
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "check_sgemm_hip")
    subprocess.run([cc, "-O2", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
                    "-L", LIBDIR, "-lugemm_hip", f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", exe], check=True)
    return exe


def test_harness_compiles_and_links_against_the_c_abi(tmp_path):
    """Plain C (no HIP headers) compiles against include/ugemm_hip.h and links
    the library; without a GPU the run stops at ugemm_hip_init with a status."""
    exe = _build(tmp_path)
    import torch
    if torch.cuda.is_available():
        return
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ugemm_hip_init" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["M=0"], "Bad parameter 'M=0'"),
    (["alpha=x"], "Not a number"),
    (["K=64", "lda=10"], "Bad parameter lda=10"),
    (["major=R", "transa=T", "M=50", "lda=40"], "Bad parameter lda=40"),   # 'T' A: lda >= M
    (["mode=fp64"], "Unknown mode"),
    (["Q=3"], "Unknown key"),
    (["layout=RXN"], "layout is"),
])
def test_harness_rejects_bad_parameters(tmp_path, args, msg):
    """The key=value command line of check_sgemm.c:145-237: a bad value or an
    ld below the stored row length exits 1 before any device work."""
    exe = _build(tmp_path)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and msg in r.stderr, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    [],                                                   # 128 x 361 x 1152 R NN, alpha 1, beta 0, 11 inputs
    ["M=77", "N=45", "K=123", "alpha=1.5", "beta=0.5", "major=C", "transa=T", "iter=3"],
    ["M=64", "N=100", "K=33", "alpha=-1", "beta=0", "transb=T", "ldb=40", "ldc=120", "iter=3"],
    ["M=31", "N=17", "K=250", "alpha=0.75", "beta=1", "major=C", "transa=T", "transb=T", "iter=3"],
    ["M=40", "N=9", "K=1", "alpha=0", "beta=0.5", "transa=T", "iter=3"],   # alpha == 0: C scaled by beta only
    ["M=200", "N=150", "K=300", "alpha=1.5", "beta=0.5", "mode=fp16x3", "seed=7", "iter=3"],
    ["M=130", "N=70", "K=260", "mode=bf16x3", "layout=RTN", "seed=0x1234", "iter=3"],
    ["M=129", "N=140", "K=256", "mode=fp8", "iter=2"],
    ["M=96", "N=64", "K=512", "mode=fp32_mfma", "major=C", "transb=T", "iter=2"],
])
def test_harness_runs_clean(tmp_path, args):
    exe = _build(tmp_path)
    print("check_sgemm_hip", *args, flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.rstrip().endswith("OK")
    # the reference's cmp_results line for every input
    iters = next((int(a.split("=")[1]) for a in args if a.startswith("iter=")), 11)
    lines = [l for l in r.stdout.splitlines() if "/" in l and " at [" in l]
    assert len(lines) == iters     # (its FAIL/Sucks verdict is informational: pass/fail is the bound)
