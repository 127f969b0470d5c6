"""Parity of the FP32-MFMA form (UGEMM_MODE_FP32_MFMA: FP32 operands, FP32
products and accumulation on the matrix cores) with the CPU oracle, to the
plain FP32 bound: |C - C_exact| <= 2(K+1) eps32 (|alpha| |A||B| + |beta| |C0|)
elementwise, C_exact from the oracle's compensated FP64 evaluation of the
reference's semantics (ugemm.h:287-411).  Small-integer operands pin the
index arithmetic bit for bit (every partial sum below 2^24 is exact)."""
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case

pytestmark = pytest.mark.gpu

MODE = ug.MODE_FP32_MFMA
COMBOS = [(m, a, b) for m in "RC" for a in "NT" for b in "NT"]
SHAPES = [(1, 1, 1), (7, 5, 9), (17, 17, 17), (33, 65, 31), (64, 64, 64), (129, 127, 65), (200, 300, 257)]
AB = [(1.0, 0.0), (1.5, 0.5), (-0.75, 1.0)]


def run_host(case, mode=MODE):
    C = case.C0.copy()
    ug.sgemm(case.major, case.ta, case.tb, case.M, case.N, case.K, case.alpha, case.A, case.lda, case.B, case.ldb,
             case.beta, C, case.ldc, mode=mode)
    return C


@pytest.mark.parametrize("major,ta,tb", COMBOS)
@pytest.mark.parametrize("M,N,K", SHAPES)
def test_layouts_shapes(major, ta, tb, M, N, K):
    for i, (alpha, beta) in enumerate(AB):
        case = Case(major, ta, tb, M, N, K, alpha, beta,
                    seed=zlib.crc32(f"m32{major}{ta}{tb}{M},{N},{K},{i}".encode()))
        ok, worst = case.judge(run_host(case), mode=MODE)
        assert ok, f"{major}{ta}{tb} {M}x{N}x{K} a={alpha} b={beta}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("variant", range(6))
@pytest.mark.parametrize("major,ta,tb", COMBOS)
def test_integer_exact_every_variant(variant, major, ta, tb):
    assert len(ug.variants(MODE)) == 6
    M, N, K = 261, 197, 88
    case = Case(major, ta, tb, M, N, K, 1.0, 1.0, seed=7)
    g = orc.rng(8 + variant)
    case.A = g.integers(-4, 5, size=case.A.shape).astype(np.float32)
    case.B = g.integers(-4, 5, size=case.B.shape).astype(np.float32)
    case.C0 = g.integers(-64, 65, size=case.C0.shape).astype(np.float32)
    ug.force_variant(MODE, variant)
    try:
        C = run_host(case)
    finally:
        ug.force_variant(MODE, -1)
    R = case.C0.copy()
    orc.sgemm_naive(major, ta, tb, M, N, K, 1.0, case.A, case.lda, case.B, case.ldb, 1.0, R, case.ldc)
    assert np.array_equal(C, R), f"variant {variant} {major}{ta}{tb}"


@pytest.mark.parametrize("variant", range(6))
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_every_variant_bound(variant, ta, tb):
    ug.force_variant(MODE, variant)
    try:
        for (M, N, K) in ((130, 250, 70), (256, 256, 48), (520, 600, 33), (300, 280, 1000)):
            case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=31 + variant)
            ok, worst = case.judge(run_host(case), mode=MODE)
            assert ok, f"variant {variant} {M}x{N}x{K}: {worst:.3g}"
    finally:
        ug.force_variant(MODE, -1)


@pytest.mark.parametrize("major,ta,tb", COMBOS)
@pytest.mark.parametrize("pad", [1, 3, 4])
def test_leading_dimensions(major, ta, tb, pad):
    case = Case(major, ta, tb, 131, 77, 93, 1.5, 0.5, seed=11 + pad, pad=pad)
    C = run_host(case)
    ok, worst = case.judge(C, mode=MODE)
    assert ok, f"worst/bound {worst:.3g}"
    assert case.untouched_padding(C)


def test_beta_zero_never_reads_c_and_alpha_zero():
    case = Case("R", "N", "N", 70, 90, 50, 1.0, 0.0, seed=3)
    case.C0[:] = np.nan
    C = run_host(case)
    assert np.all(np.isfinite(C))
    case.C0[:] = 0.0
    assert case.judge(C, mode=MODE)[0]
    case = Case("R", "N", "T", 37, 29, 11, 0.0, 0.5, seed=5, pad=2)
    C = run_host(case)
    assert np.array_equal(C[:, :case.N], case.C0[:, :case.N] * np.float32(0.5))
    assert case.untouched_padding(C)


def test_misaligned_bases_use_scalar_loads():
    M, N, K = 300, 260, 200
    case = Case("R", "N", "N", M, N, K, 1.5, 0.5, seed=41)
    s = hip.Stream()
    for off in (0, 1):
        dA = hip.DeviceArray(case.A.size + 1, np.float32)
        dB = hip.DeviceArray(case.B.size + 1, np.float32)
        dC = hip.DeviceArray(case.C0.size + 1, np.float32)
        hA = np.zeros(case.A.size + 1, np.float32); hA[off:off + case.A.size] = case.A.ravel()
        hB = np.zeros(case.B.size + 1, np.float32); hB[off:off + case.B.size] = case.B.ravel()
        hC = np.zeros(case.C0.size + 1, np.float32); hC[off:off + case.C0.size] = case.C0.ravel()
        dA.upload(hA); dB.upload(hB); dC.upload(hC)
        ug.sgemm_dev("R", "N", "N", M, N, K, case.alpha, dA.at(off), case.lda, dB.at(off), case.ldb, case.beta,
                     dC.at(off), case.ldc, mode=MODE, stream=s)
        s.synchronize()
        C = dC.download()[off:off + case.C0.size].reshape(case.C0.shape)
        ok, worst = case.judge(C, mode=MODE)
        assert ok, f"offset {off}: {worst:.3g}"
    s.destroy()


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N")])
def test_large_sampled(ta, tb):
    M = N = K = 2048
    case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=51)
    C = run_host(case)
    rows = np.sort(orc.rng(52).choice(M, 64, replace=False))
    exact, scale = case.exact(rows=rows)
    ok, worst = orc.check_bound(C[rows, :N], exact, scale, orc.tolerance(K))
    assert ok, f"worst/bound {worst:.3g}"


@pytest.mark.parametrize("splits", [2, 5])
def test_split_k_and_determinism(splits):
    ug.force_splits(MODE, splits)
    try:
        for ta, tb in (("N", "N"), ("T", "N"), ("N", "T")):
            case = Case("R", ta, tb, 150, 140, 4100, 1.5, 0.5, seed=splits + len(ta + tb))
            C1, C2 = run_host(case), run_host(case)
            assert np.array_equal(C1.view(np.uint32), C2.view(np.uint32))
            ok, worst = case.judge(C1, mode=MODE)
            assert ok, f"splits {splits} {ta}{tb}: {worst:.3g}"
    finally:
        ug.force_splits(MODE, 0)


@pytest.mark.parametrize("splits", [0, 3])
def test_bias_leaky_relu_epilogue(splits):
    M, N, K = 96, 200, 700
    g = orc.rng(61 + splits)
    A, B = orc.uniform(g, (M, K)), orc.uniform(g, (K, N))
    C0, bias = orc.uniform(g, (M, N)), orc.uniform(g, (N,))
    dA, dB, dC, db = (hip.DeviceArray.from_host(x) for x in (A, B, C0, bias))
    ug.force_splits(MODE, splits)
    try:
        ug.sgemm_ex_dev("R", "N", "N", M, N, K, 1.25, dA, K, dB, N, 0.5, dC, N, mode=MODE, bias=db,
                        bias_dim=ug.BIAS_COLS, act=ug.ACT_LEAKY_RELU, slope=0.1)
        hip.synchronize()
    finally:
        ug.force_splits(MODE, 0)
    got = dC.download().astype(np.float64)
    pre = 1.25 * (A.astype(np.float64) @ B.astype(np.float64)) + 0.5 * C0 + bias[None, :]
    want = np.where(pre >= 0, pre, 0.1 * pre)
    scale = 1.25 * (np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))) + 0.5 * np.abs(C0) + \
        np.abs(bias[None, :])
    assert np.all(np.abs(got - want) <= orc.tolerance(K) * scale + 1e-30)


def test_conv_argument_checks():
    L = ug.load()
    assert ug.MODES["fp32_mfma"] == MODE
    # NULL buffers are refused before any kernel is launched.
    assert L.ugemm_hip_conv2d_dev(None, 1, 4, 4, None, 3, 1, 1, None, 1, None, 0, 0.1, MODE, None) == ug.ERR_ARG


def test_tuner_and_every_variant_named():
    names = ug.variants(MODE)
    assert len(names) == 6 and all(n.startswith("f32_mfma_") for n in names)
    L = ug.load()
    v = L.ugemm_hip_tune_sgemm(b"N", b"N", 1024, 768, 512, MODE)
    assert 0 <= v < 6
    assert L.ugemm_hip_tune_lookup(b"N", b"N", 1024, 768, 512, MODE) == v
    case = Case("R", "N", "N", 1024, 768, 512, 1.0, 0.0, seed=71)
    assert case.judge(run_host(case), mode=MODE)[0]
