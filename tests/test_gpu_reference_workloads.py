"""The reference's own GPU-backend workload and the edge sizes of its test
strategy, through the drop-in entry points.

* sgemm_test.c: M = 1023, N = 1000, K = 1023, A[i] = B[i] = i + 1, C = 1,
  alpha = 1, beta = 0, in the RNN, RNT and RTN layouts (sgemm_test.c:168-172,
  :219-222, :233-258), through sgemm_hip_packed -- the packed sgemm_ocl
  argument list that sgemm_test.c's macro block binds -- where the reference
  compares with host loops; here every row is compared with the FP64 oracle
  (the values reach 1e15, so this also exercises the bound's scale), and the
  reference's cmp_results line is printed.
* Edge sizes 63, 65 and 1023 (one below and above a 64-wide tile, and the
  sgemm_test.c size) in all 8 major/transpose combinations, for the plain
  FP32, FP32-MFMA and the FP16X3 / BF16X3 forms."""
import itertools
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from oracle import oracle as orc
from tests.helpers import Case, MODE_EXTRA

pytestmark = pytest.mark.gpu


def sgemm_test_operands():
    M, N, K = 1023, 1000, 1023
    A = np.arange(1, M * K + 1, dtype=np.float64).astype(np.float32)   # exact: < 2^24
    B = np.arange(1, K * N + 1, dtype=np.float64).astype(np.float32)
    return M, N, K, A, B


_EXACT = {}   # FP64 product and error scale per layout (the same for every form)


@pytest.mark.parametrize("layout", ["RNN", "RNT", "RTN"])
@pytest.mark.parametrize("mode", [None, ug.MODE_FP16X3, ug.MODE_FP32_MFMA])
def test_sgemm_test_workload(layout, mode):
    M, N, K, A, B = sgemm_test_operands()
    ta, tb = layout[1], layout[2]
    lda, ldb = (M if ta == "T" else K), (K if tb == "T" else N)
    C = np.ones(M * N, np.float32)
    print(f"sgemm_test.c workload {layout} mode {mode}", flush=True)
    if mode is None:
        L = ug.load()
        L.sgemm_hip_packed(ta.encode(), tb.encode(), M, N, K, 1.0, A.ctypes.data, B.ctypes.data, 0.0, C.ctypes.data)
        assert L.ugemm_hip_last_error() == ug.OK, L.ugemm_hip_last_error_message()
    else:
        ug.sgemm("R", ta, tb, M, N, K, 1.0, A, lda, B, ldb, 0.0, C, N, mode=mode)
    if layout not in _EXACT:
        _EXACT[layout] = orc.reference_f64("R", ta, tb, M, N, K, 1.0, A, lda, B, ldb, 0.0, None, N)
    exact, scale = _EXACT[layout]
    std_err, std_ref, max_err, (mi, ni), verdict = orc.cmp_results(M, N, exact.astype(np.float32), C, N)
    print(f"{std_err:.3e}/{std_ref:.3e}={std_err / std_ref:.3e}. {max_err:.3e} at [{mi:3d},{ni:3d}] {verdict}")
    ok, worst = orc.check_bound(C.reshape(M, N), exact, scale, orc.tolerance(K, extra=MODE_EXTRA.get(mode or 0, 0.0)))
    assert ok, f"{layout} mode {mode}: worst/bound {worst:.3g}"


EDGE_SHAPES = sorted(set(itertools.permutations((63, 65, 1023))))


@pytest.mark.parametrize("mode", [ug.MODE_FP32, ug.MODE_FP32_MFMA, ug.MODE_FP16X3, ug.MODE_BF16X3])
@pytest.mark.parametrize("major,ta,tb", [(m, a, b) for m in "RC" for a in "NT" for b in "NT"])
def test_edge_sizes(mode, major, ta, tb):
    for (M, N, K) in EDGE_SHAPES:
        seed = zlib.crc32(f"edge{mode}{major}{ta}{tb}{M},{N},{K}".encode())
        case = Case(major, ta, tb, M, N, K, 1.5, 0.5, seed=seed, pad=1)
        C = case.C0.copy()
        ug.sgemm(major, ta, tb, M, N, K, 1.5, case.A, case.lda, case.B, case.ldb, 0.5, C, case.ldc, mode=mode)
        ok, worst = case.judge(C, mode=mode)
        assert ok, f"mode {mode} {major}{ta}{tb} {M}x{N}x{K}: worst/bound {worst:.3g}"
        assert case.untouched_padding(C)
