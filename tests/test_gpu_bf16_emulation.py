"""BF16X1, BF16X3 and BF16X6 element by element against the exact emulation
of their split (oracle.bf16_product: round-to-nearest-even bf16 planes and
the product set the split core issues).

The GPU result and the emulation differ only by FP32 accumulation, so the
plain FP32 bound applies.  At the small K used first, that bound is tight
enough to reject a split that truncates instead of rounding
(tests/test_oracle.py plants that mistake in the oracle); the larger K and
the 2048^3 cases check the same emulation at size on sampled rows."""
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case

pytestmark = pytest.mark.gpu

FORMS = [(ug.MODE_BF16X1, 1), (ug.MODE_BF16X3, 2), (ug.MODE_BF16X6, 3)]
COMBOS = [(m, a, b) for m in "RC" for a in "NT" for b in "NT"]


def emulation_check(case, C, NP, rows=None):
    """(ok, worst) of C (the logical M x N result, alpha = 1, beta = 0) against
    the emulation, on `rows` of C (all when None)."""
    Aop = orc.op_matrix(case.major, case.ta, case.A, case.M, case.K)
    Bop = orc.op_matrix(case.major, case.tb, case.B, case.K, case.N)
    if rows is not None:
        Aop = Aop[rows]
    emu = orc.bf16_product(Aop, Bop, NP)
    scale = np.abs(Aop.astype(np.float64)) @ np.abs(Bop.astype(np.float64))
    got = case.logical(C) if rows is None else case.logical(C)[rows]
    return orc.check_bound(got, emu, scale, orc.tolerance(case.K))


@pytest.mark.parametrize("mode,NP", FORMS, ids=["bf16x1", "bf16x3", "bf16x6"])
@pytest.mark.parametrize("major,ta,tb", COMBOS)
def test_bf16_forms_match_emulation(mode, NP, major, ta, tb):
    # K = 1 and 5: the bound is tight enough to tell rounding from truncation
    # (BF16X6 only at K = 1 on 256 x 256); K = 100 and 300 cross several
    # 32-wide K steps and a padded tail.
    shapes = [(256, 256, 1)] if NP == 3 else [(64, 64, 5), (70, 90, 1)]
    shapes += [(96, 80, 100), (33, 65, 300)]
    for (M, N, K) in shapes:
        seed = zlib.crc32(f"bfemu{mode}{major}{ta}{tb}{M},{N},{K}".encode())
        case = Case(major, ta, tb, M, N, K, 1.0, 0.0, seed=seed)
        C = case.C0.copy()
        ug.sgemm(major, ta, tb, M, N, K, 1.0, case.A, case.lda, case.B, case.ldb, 0.0, C, case.ldc, mode=mode)
        ok, worst = emulation_check(case, C, NP)
        assert ok, f"mode {mode} {major}{ta}{tb} {M}x{N}x{K}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("mode,NP", FORMS, ids=["bf16x1", "bf16x3", "bf16x6"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N")])
def test_bf16_forms_match_emulation_2048(mode, NP, ta, tb):
    """2048^3 on device buffers, 24 sampled rows of C against the emulation."""
    n = 2048
    case = Case("R", ta, tb, n, n, n, 1.0, 0.0, seed=zlib.crc32(f"bfemu2048{mode}{ta}{tb}".encode()))
    print(f"bf16 emulation 2048^3 mode {mode} {ta}{tb}", flush=True)
    dA, dB = hip.DeviceArray.from_host(case.A), hip.DeviceArray.from_host(case.B)
    dC = hip.DeviceArray((n, n), np.float32)
    try:
        ug.sgemm_dev("R", ta, tb, n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n, mode=mode)
        C = dC.download()
    finally:
        for d in (dA, dB, dC):
            d.free()
    rows = np.sort(np.random.Generator(np.random.PCG64(11)).choice(n, 24, replace=False))
    ok, worst = emulation_check(case, C, NP, rows=rows)
    assert ok, f"mode {mode} {ta}{tb} 2048^3: worst/bound {worst:.3g}"
