"""FP8-e4m3 mode (one power-of-two scale per 128x128 block): the GPU result
must equal the oracle's emulation -- the same block quantisation, then an
exact product -- up to accumulation error,
|C - C_emul| <= (2(K+1) eps32 + 2^-12) * (|alpha| |Aq||Bq| + |beta| |C0|).
The 2^-12 term covers the 16x16x128 FP8 MFMA's internal summation of its
128 products, measured at ~2^-16 of sum |a b| on gfx950 (below FP32
sequential accuracy); it is 250x smaller than one e4m3 rounding step (2^-4),
so any quantisation mismatch fails.  There is no FP8 path in the reference."""
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from oracle import oracle as orc
from tests.helpers import Case

FP8_EXTRA = 2.0 ** -12

pytestmark = pytest.mark.gpu


def emulate(case):
    A = orc.op_matrix(case.major, case.ta, case.A, case.M, case.K)
    B = orc.op_matrix(case.major, case.tb, case.B, case.K, case.N)
    Aq = orc.fp8_block_dequant(A)
    Bq = orc.fp8_block_dequant(B.T).T          # B blocks are (n, k) tiles
    C0 = orc.c_matrix(case.major, case.C0, case.M, case.N).astype(np.float64)
    exact = case.alpha * (Aq @ Bq) + case.beta * C0
    scale = abs(case.alpha) * (np.abs(Aq) @ np.abs(Bq)) + abs(case.beta) * np.abs(C0)
    return exact, scale


def run(case):
    C = case.C0.copy()
    ug.sgemm(case.major, case.ta, case.tb, case.M, case.N, case.K, case.alpha, case.A, case.lda, case.B, case.ldb,
             case.beta, C, case.ldc, mode=ug.MODE_FP8_BLOCK)
    return C


@pytest.mark.parametrize("major,ta,tb", [(m, a, b) for m in "RC" for a in "NT" for b in "NT"])
def test_fp8_matches_emulation(major, ta, tb):
    for (M, N, K) in ((1, 1, 1), (33, 65, 31), (200, 300, 257), (129, 260, 384)):
        seed = zlib.crc32(f"f8{major}{ta}{tb}{M}{N}{K}".encode())
        case = Case(major, ta, tb, M, N, K, 1.5, 0.5, seed=seed)
        C = run(case)
        exact, scale = emulate(case)
        ok, worst = orc.check_bound(case.logical(C), exact, scale, orc.tolerance(K, extra=FP8_EXTRA))
        assert ok, f"{major}{ta}{tb} {M}x{N}x{K}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_fp8_variants_and_wide_range(variant):
    """Every compiled tile variant against the emulation, including the
    8-wave interleaved one at 1, 2 and 5 K tiles (its pipeline's edges)."""
    assert len(ug.variants(ug.MODE_FP8_BLOCK)) == 3
    ug.force_variant(ug.MODE_FP8_BLOCK, variant)
    try:
        for (M, N, K, beta, ta, tb) in ((300, 280, 520, 0.0, "N", "N"), (600, 530, 128, 0.5, "N", "T"),
                                        (257, 513, 256, 0.5, "T", "N"), (520, 300, 700, 1.0, "T", "T")):
            case = Case("R", ta, tb, M, N, K, 1.0, beta, seed=17 + variant + M, dist="wide")
            C = run(case)
            exact, scale = emulate(case)
            ok, worst = orc.check_bound(case.logical(C), exact, scale, orc.tolerance(case.K, extra=FP8_EXTRA))
            assert ok, f"variant {variant} {ta}{tb} {M}x{N}x{K}: {worst:.3g}"
    finally:
        ug.force_variant(ug.MODE_FP8_BLOCK, -1)


def test_fp8_zero_blocks_and_exact_values():
    # Small integers are exact in e4m3 (|v| <= 16) and the block scale of a
    # block with amax 16 is 2^-4 ... the products must be exact.
    M, N, K = 130, 140, 256
    g = orc.rng(3)
    A = g.integers(-16, 17, size=(M, K)).astype(np.float32)
    B = g.integers(-16, 17, size=(K, N)).astype(np.float32)
    A[:, :128] = 0.0                                  # whole zero blocks
    C = np.zeros((M, N), np.float32)
    ug.sgemm("R", "N", "N", M, N, K, 1.0, A, K, B, N, 0.0, C, N, mode=ug.MODE_FP8_BLOCK)
    assert np.array_equal(C, (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32))
