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


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    [],                                                   # 128 x 361 x 1152 R NN (the harness default shape)
    ["77", "45", "123", "1.5", "0.5", "C", "T", "N"],
    ["64", "100", "33", "-1", "0", "R", "N", "T"],
    ["31", "17", "250", "0.75", "1", "C", "T", "T"],
    ["40", "9", "1", "0", "0.5", "R", "T", "N"],          # alpha == 0: C scaled by beta only
])
def test_harness_runs_clean(tmp_path, args):
    exe = _build(tmp_path)
    print("check_sgemm_hip", *args, flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.rstrip().endswith("OK")
