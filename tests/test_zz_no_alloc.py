"""ugemm_hip_sgemm_noalloc_dev: the SGEMM whose workspace comes from the
library's static device scratch instead of the allocator.

Each case makes two calls in a row through the C ABI -- the second on a
smaller input, so that stale scratch contents of the first (planes, FP16X3
row maxima, split-K partial products) would show -- and compares each with
ugemm_hip_sgemm_dev on the same inputs: the same kernels run, so the results
must be bitwise equal; non-FP8 forms are also held to their oracle bound.
A call whose workspace exceeds the scratch must fail before touching C."""
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case

pytestmark = pytest.mark.gpu

# (mode, first call, second call): (ta, tb, M, N, K) each; the first shapes
# need workspace (split planes, FP8 planes, or split-K partial products for
# the FP32 forms: few tiles and a long K).
CASES = [
    (ug.MODE_FP32, ("N", "N", 256, 256, 8192), ("T", "N", 96, 80, 3000)),
    (ug.MODE_FP32_MFMA, ("N", "T", 256, 192, 8192), ("N", "N", 64, 100, 2100)),
    (ug.MODE_BF16X3, ("N", "N", 1024, 768, 2048), ("T", "N", 300, 200, 500)),
    (ug.MODE_FP16X3, ("N", "N", 1024, 768, 2048), ("N", "T", 300, 200, 500)),
    (ug.MODE_BF16X1, ("T", "T", 768, 512, 1024), ("N", "N", 130, 70, 90)),
    (ug.MODE_BF16X6, ("N", "T", 512, 512, 1536), ("T", "N", 200, 300, 400)),
    (ug.MODE_FP8_BLOCK, ("N", "N", 1024, 1024, 1024), ("T", "N", 300, 257, 500)),
]


def run_both(case, mode):
    """(result of the no-allocation entry, result of ugemm_hip_sgemm_dev)."""
    dA, dB = hip.DeviceArray.from_host(case.A), hip.DeviceArray.from_host(case.B)
    dC1, dC2 = hip.DeviceArray.from_host(case.C0), hip.DeviceArray.from_host(case.C0)
    try:
        args = (case.major, case.ta, case.tb, case.M, case.N, case.K, case.alpha, dA, case.lda, dB, case.ldb,
                case.beta)
        ug.sgemm_noalloc_dev(*args, dC1, case.ldc, mode=mode)
        ug.sgemm_dev(*args, dC2, case.ldc, mode=mode)
        hip.synchronize()
        return dC1.download(), dC2.download()
    finally:
        for d in (dA, dB, dC1, dC2):
            d.free()


@pytest.mark.parametrize("mode,first,second", CASES, ids=[f"mode{c[0]}" for c in CASES])
def test_noalloc_matches_dev_entry(mode, first, second):
    for call, (ta, tb, M, N, K) in enumerate((first, second)):
        seed = zlib.crc32(f"noalloc{mode}{call}{ta}{tb}{M},{N},{K}".encode())
        case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=seed)
        print(f"noalloc mode {mode} call {call}: {ta}{tb} {M}x{N}x{K}", flush=True)
        got, ref = run_both(case, mode)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
            f"mode {mode} call {call}: differs from ugemm_hip_sgemm_dev"
        if mode != ug.MODE_FP8_BLOCK:
            ok, worst = case.judge(got, mode=mode)
            assert ok, f"mode {mode} call {call}: worst/bound {worst:.3g}"


def test_noalloc_column_major_and_alpha_zero():
    """Column-major normalisation and the alpha == 0 path (no workspace)."""
    for (major, alpha, mode) in (("C", 1.5, ug.MODE_FP16X3), ("R", 0.0, ug.MODE_BF16X3)):
        case = Case(major, "N", "T", 200, 150, 700, alpha, 0.5, seed=zlib.crc32(f"na{major}{alpha}".encode()))
        got, ref = run_both(case, mode)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        ok, worst = case.judge(got, mode=mode)
        assert ok, worst


def test_noalloc_rejects_workspace_beyond_scratch():
    """BF16x3 at 4096^3 needs 2 planes x 2 B x (4096 + 4096) x 4096 = 128 MiB of
    planes: more than the scratch. The call fails with ERR_SCRATCH before any
    launch, and C keeps its contents; a fitting call right after still works."""
    n = 4096
    need = 2 * 2 * (n + n) * n
    assert need > ug.NOALLOC_SCRATCH_BYTES
    dA = hip.DeviceArray((n, n), np.float32)
    dB = hip.DeviceArray((n, n), np.float32)
    sentinel = np.full((n, n), 7.0, np.float32)
    dC = hip.DeviceArray.from_host(sentinel)
    try:
        with pytest.raises(ug.UgemmError) as e:
            ug.sgemm_noalloc_dev("R", "N", "N", n, n, n, 1.0, dA, n, dB, n, 0.5, dC, n, mode=ug.MODE_BF16X3)
        assert e.value.status == ug.ERR_SCRATCH
        assert "static scratch" in str(e.value)
        hip.synchronize()
        assert np.array_equal(dC.download(), sentinel)
    finally:
        for d in (dA, dB, dC):
            d.free()
    case = Case("R", "N", "N", 256, 256, 512, 1.0, 0.0, seed=5)
    got, ref = run_both(case, ug.MODE_BF16X3)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_noalloc_status_code_string():
    L = ug.load()
    assert b"static scratch" in L.ugemm_hip_status_string(ug.ERR_SCRATCH)
