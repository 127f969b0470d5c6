"""Parity of the split MFMA forms (BF16X1, BF16X3, BF16X6, FP16X3) with the oracle.

Acceptance per element: |C - C_exact| <= (2(K+1) eps32 + extra) * scale with
extra = 2^-6 (x1), 2^-14 (x3), 2^-21 (x6), 2^-19 (fp16x3) -- the per-product
error of the dropped split terms (see tests/helpers.py).  FP16X3 is also
compared with its exact emulation (oracle.fp16x3_product), where only the
FP32 accumulation may differ."""
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case

pytestmark = pytest.mark.gpu

MODES = [ug.MODE_BF16X1, ug.MODE_BF16X3, ug.MODE_BF16X6, ug.MODE_FP16X3]
COMBOS = [(m, a, b) for m in "RC" for a in "NT" for b in "NT"]
SHAPES = [(1, 1, 1), (7, 5, 9), (33, 65, 31), (129, 127, 65), (200, 300, 257)]


def run_host(case, mode):
    C = case.C0.copy()
    ug.sgemm(case.major, case.ta, case.tb, case.M, case.N, case.K, case.alpha, case.A, case.lda, case.B, case.ldb,
             case.beta, C, case.ldc, mode=mode)
    return C


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("major,ta,tb", COMBOS)
def test_split_layouts_shapes(mode, major, ta, tb):
    for (M, N, K) in SHAPES:
        seed = zlib.crc32(f"{mode}{major}{ta}{tb}{M},{N},{K}".encode())
        case = Case(major, ta, tb, M, N, K, 1.5, 0.5, seed=seed)
        C = run_host(case, mode)
        ok, worst = case.judge(C, mode=mode)
        assert ok, f"mode {mode} {major}{ta}{tb} {M}x{N}x{K}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("variant", range(8))
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_split_every_variant(mode, variant, ta, tb):
    assert len(ug.variants(mode)) == 8
    ug.force_variant(mode, variant)
    try:
        # The last three have three or more 256-wide tiles on both sides (no
        # operand is split in the core) and 4, 1 and 2 K tiles of 32: the
        # streamed staging of variants 6 and 7 at its pipeline's edges.
        for (M, N, K) in ((130, 250, 70), (300, 260, 96), (600, 700, 100), (520, 530, 20), (530, 600, 40)):
            case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=31 + variant)
            C = run_host(case, mode)
            ok, worst = case.judge(C, mode=mode)
            assert ok, f"variant {variant} {M}x{N}x{K}: {worst:.3g}"
    finally:
        ug.force_variant(mode, -1)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape,splits", [((1000, 1100, 1200, "N", "N"), 1), ((777, 513, 333, "T", "T"), 1),
                                          ((600, 530, 4096, "T", "N"), 3), ((530, 600, 40, "N", "T"), 1)])
def test_split_variants_bitwise_identical(mode, shape, splits):
    """The variants differ in tiles, waves and the order of their LDS reads,
    never in the MFMAs an accumulator receives or their order: at one split
    count, every variant gives the same bits."""
    M, N, K, ta, tb = shape
    case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=zlib.crc32(f"bitwise{mode}{shape}".encode()))
    outs = []
    for v in range(len(ug.variants(mode))):
        ug.force_variant(mode, v)
        ug.force_splits(mode, splits)   # the same split count for every variant
        try:
            outs.append(run_host(case, mode))
        finally:
            ug.force_variant(mode, -1)
            ug.force_splits(mode, 0)
    ok, worst = case.judge(outs[0], mode=mode)
    assert ok, f"{worst:.3g}"
    for v in range(1, len(outs)):
        assert np.array_equal(outs[0].view(np.uint32), outs[v].view(np.uint32)), f"variant {v} differs from 0"


@pytest.mark.parametrize("mode", [ug.MODE_BF16X3, ug.MODE_FP16X3, ug.MODE_BF16X1, ug.MODE_BF16X6])
@pytest.mark.parametrize("variant", [4, 6, 7])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("pad", [0, 1])
def test_split_operand_fused_in_core(mode, variant, ta, tb, pad):
    """The 8-wave variants split an operand that at most two tiles re-read
    while staging it, from either FP32 layout: k-contiguous (A in N, B in T)
    or row-contiguous (A in T, B in N), the latter transposed in registers
    (two rows per thread with 256-row tiles, one with 128-row tiles: the
    128 x 128 variant 7 and the 256 x 128 tiles of BF16x6).  One side of
    each shape has at most 2 tiles of 256 (the last two: of 128), the
    other 5 or more; odd row counts and pad = 1 (odd leading dimensions:
    scalar loads) hit the edges."""
    ug.force_variant(mode, variant)
    try:
        for (M, N, K) in ((301, 1100, 77), (1100, 257, 96), (3, 1031, 40), (1030, 511, 33), (1100, 200, 77),
                          (250, 1031, 96)):
            case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=M * 3 + N + K + pad, pad=pad)
            C = run_host(case, mode)
            ok, worst = case.judge(C, mode=mode)
            assert ok, f"variant {variant} {ta}{tb} {M}x{N}x{K} pad {pad}: {worst:.3g}"
    finally:
        ug.force_variant(mode, -1)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("pad", [1, 4])
def test_split_leading_dimensions_and_nan_c(mode, pad):
    case = Case("R", "T", "T", 131, 77, 93, 1.5, 0.5, seed=11 + pad, pad=pad)
    C = run_host(case, mode)
    assert case.judge(C, mode=mode)[0]
    assert case.untouched_padding(C)
    case = Case("R", "N", "N", 70, 90, 50, 1.0, 0.0, seed=3)
    case.C0[:] = np.nan
    C = run_host(case, mode)
    assert np.all(np.isfinite(C))


def test_split_accuracy_ordering():
    # K = 4096, U[-1,1]: rms error vs exact -- bf16x6 and fp16x3 at or below
    # plain FP32, bf16x3 within a small factor of it, bf16x1 far above (8-bit
    # operands).
    M = N = 256
    K = 4096
    # One K split for every form: this small grid would otherwise be split
    # along K, whose shorter sequential sums make plain FP32 several times
    # more accurate (its error is accumulation error; bf16x3's is the
    # dropped lo*lo term), and the comparison is of the unsplit kernels.
    case = Case("R", "N", "N", M, N, K, 1.0, 0.0, seed=77)
    exact, _ = case.exact()
    rms = {}
    for mode in (ug.MODE_FP32, ug.MODE_BF16X1, ug.MODE_BF16X3, ug.MODE_BF16X6, ug.MODE_FP16X3):
        ug.force_splits(mode, 1)
        try:
            C = run_host(case, mode)
        finally:
            ug.force_splits(mode, 0)
        rms[mode] = float(np.sqrt(np.mean((C.astype(np.float64) - exact) ** 2)))
    assert rms[ug.MODE_BF16X6] <= rms[ug.MODE_FP32], rms
    assert rms[ug.MODE_FP16X3] <= rms[ug.MODE_FP32], rms
    assert rms[ug.MODE_BF16X3] <= 4 * rms[ug.MODE_FP32], rms
    assert rms[ug.MODE_BF16X1] > 100 * rms[ug.MODE_FP32], rms


def test_split_special_values():
    # An infinite operand: BF16X1 keeps +inf; the split forms multiply the
    # infinite leading plane by the other operand's trailing planes, which are
    # 0 for bf16-exact values, so inf * 0 makes those entries NaN (documented
    # behaviour of split GEMMs). Rows without the inf stay exact.
    M, N, K = 16, 16, 32
    A = np.ones((M, K), np.float32)
    B = np.ones((K, N), np.float32)
    A[3, 5] = np.inf
    for mode in MODES:
        C = np.zeros((M, N), np.float32)
        ug.sgemm("R", "N", "N", M, N, K, 1.0, A, K, B, N, 0.0, C, N, mode=mode)
        if mode == ug.MODE_BF16X1:
            assert np.all(np.isposinf(C[3])), mode
        else:
            assert not np.any(np.isfinite(C[3])), mode
        assert np.all(C[np.arange(M) != 3] == K), mode


@pytest.mark.parametrize("mode", MODES)
def test_split_large_sampled(mode):
    M = N = K = 2048
    case = Case("R", "N", "N", M, N, K, 1.5, 0.5, seed=51)
    C = run_host(case, mode)
    rows = np.sort(orc.rng(52).choice(M, 48, replace=False))
    exact, scale = case.exact(rows=rows)
    from tests.helpers import MODE_EXTRA
    ok, worst = orc.check_bound(C[rows, :N], exact, scale, orc.tolerance(K, extra=MODE_EXTRA[mode]))
    assert ok, f"worst/bound {worst:.3g}"


@pytest.mark.parametrize("dist", ["uniform", "normal", "wide"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_fp16x3_matches_emulation(dist, ta, tb):
    """GPU FP16X3 vs the exact emulation of its split and scales: the two
    differ only by FP32 accumulation, so the plain FP32 bound applies; wide
    data (entries sign*2^e*(1+u), e in [-12, 12]) exercises the row scales."""
    M, N, K = 96, 80, 300
    case = Case("R", ta, tb, M, N, K, 1.0, 0.0, seed=zlib.crc32(f"f16{dist}{ta}{tb}".encode()), dist=dist)
    C = run_host(case, ug.MODE_FP16X3)
    Aop = orc.op_matrix("R", ta, case.A, M, K)
    Bop = orc.op_matrix("R", tb, case.B, K, N)
    emu = orc.fp16x3_product(Aop, Bop)
    scale = np.abs(Aop.astype(np.float64)) @ np.abs(Bop.astype(np.float64))
    ok, worst = orc.check_bound(C[:, :N], emu, scale, orc.tolerance(K))
    assert ok, f"worst/bound {worst:.3g}"


@pytest.mark.parametrize("K", [8192, 8190, 8200])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_fp16x3_row_scaled_split_at_k_limit(K, tb):
    """The split of a k-contiguous A that computes its row maxima itself
    (two rows per block, up to 8 passes of 32 K tiles: K padded to at most
    8192; NN merges it with B's column maxima, NT runs it alone) against the
    exact emulation, at and around its K limit (8200 takes the separate
    row-maximum pass); 66 rows leave padding rows in the last pair."""
    M, N = 66, 40
    case = Case("R", "N", tb, M, N, K, 1.0, 0.0, seed=K * 7 + (tb == "T"), dist="wide")
    C = run_host(case, ug.MODE_FP16X3)
    Aop = orc.op_matrix("R", "N", case.A, M, K)
    Bop = orc.op_matrix("R", tb, case.B, K, N)
    emu = orc.fp16x3_product(Aop, Bop)
    scale = np.abs(Aop.astype(np.float64)) @ np.abs(Bop.astype(np.float64))
    ok, worst = orc.check_bound(C[:, :N], emu, scale, orc.tolerance(K))
    assert ok, f"K {K} {tb}: worst/bound {worst:.3g}"


def test_fp16x3_row_scales_extreme_rows():
    """Rows and columns of very different magnitude (2^-60 .. 2^60) and an
    all-zero row: each row's scale brings it into fp16 range independently."""
    M, N, K = 8, 6, 40
    g = orc.rng(93)
    A = orc.uniform(g, (M, K)) * np.ldexp(np.float32(1.0), np.arange(-60, 60, 15))[:, None].astype(np.float32)
    A[5] = 0.0
    B = orc.uniform(g, (K, N)) * np.ldexp(np.float32(1.0), np.array([-50, -20, 0, 10, 30, 50]))[None, :].astype(np.float32)
    A = A.astype(np.float32); B = B.astype(np.float32)
    C = np.zeros((M, N), np.float32)
    ug.sgemm("R", "N", "N", M, N, K, 1.0, A, K, B, N, 0.0, C, N, mode=ug.MODE_FP16X3)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    scale = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    ok, worst = orc.check_bound(C, exact, scale, orc.tolerance(K, extra=2.0 ** -19))
    assert ok, worst
    assert np.all(C[5] == 0.0)
