"""Parity of the SGEMM path (plain FP32 and the MFMA forms) with the CPU oracle.

Every case calls libugemm_hip.so through its C ABI (host-pointer entry
ugemm_hip_sgemm, or the device-pointer entry) and is judged elementwise:
|C - C_exact| <= (2(K+1) eps32 + extra(mode)) * (|alpha| |A||B| + |beta| |C0|),
with C_exact from the oracle's compensated FP64 evaluation of the reference's
semantics (ugemm.h:287-411)."""
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case

pytestmark = pytest.mark.gpu

COMBOS = [(m, a, b) for m in "RC" for a in "NT" for b in "NT"]
SHAPES = [(1, 1, 1), (7, 5, 9), (17, 17, 17), (33, 65, 31), (64, 64, 64), (129, 127, 65), (200, 300, 257)]
AB = [(1.0, 0.0), (1.5, 0.5), (-0.75, 1.0)]


def run_host(case, mode=ug.MODE_FP32):
    C = case.C0.copy()
    ug.sgemm(case.major, case.ta, case.tb, case.M, case.N, case.K, case.alpha, case.A, case.lda, case.B, case.ldb,
             case.beta, C, case.ldc, mode=mode)
    return C


def test_known_answer():
    # sgemm_test.c:186-200: A = B = [1..6], C = A * B for a 3x2 by 2x3 product.
    A = np.arange(1, 7, dtype=np.float32)
    B = np.arange(1, 7, dtype=np.float32)
    C = np.full(9, 7.0, np.float32)
    ug.sgemm("R", "N", "N", 3, 3, 2, 1.0, A, 2, B, 3, 0.0, C, 3)
    assert C.tolist() == [9, 12, 15, 19, 26, 33, 29, 40, 51]


@pytest.mark.parametrize("major,ta,tb", COMBOS)
@pytest.mark.parametrize("M,N,K", SHAPES)
def test_fp32_layouts_shapes(major, ta, tb, M, N, K):
    for i, (alpha, beta) in enumerate(AB):
        case = Case(major, ta, tb, M, N, K, alpha, beta, seed=zlib.crc32(f"{major}{ta}{tb}{M},{N},{K},{i}".encode()))
        C = run_host(case)
        ok, worst = case.judge(C)
        assert ok, f"{major}{ta}{tb} {M}x{N}x{K} a={alpha} b={beta}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("major,ta,tb", COMBOS)
def test_fp32_integer_exact(major, ta, tb):
    # Small-integer operands: every partial sum stays below 2^24, so FP32 is
    # exact and the result must equal the naive loop bit for bit.
    M, N, K = 97, 83, 160
    case = Case(major, ta, tb, M, N, K, 1.0, 1.0, seed=7)
    g = orc.rng(8)
    case.A = g.integers(-4, 5, size=case.A.shape).astype(np.float32)
    case.B = g.integers(-4, 5, size=case.B.shape).astype(np.float32)
    case.C0 = g.integers(-64, 65, size=case.C0.shape).astype(np.float32)
    C = run_host(case)
    R = case.C0.copy()
    orc.sgemm_naive(major, ta, tb, M, N, K, 1.0, case.A, case.lda, case.B, case.ldb, 1.0, R, case.ldc)
    assert np.array_equal(C, R)


@pytest.mark.parametrize("major,ta,tb", COMBOS)
@pytest.mark.parametrize("pad", [1, 3, 4])
def test_fp32_leading_dimensions(major, ta, tb, pad):
    # ld > row length (unaligned for pad 1 and 3): the gaps between rows of C
    # must keep their contents.
    case = Case(major, ta, tb, 131, 77, 93, 1.5, 0.5, seed=11 + pad, pad=pad)
    C = run_host(case)
    ok, worst = case.judge(C)
    assert ok, f"worst/bound {worst:.3g}"
    assert case.untouched_padding(C)


def test_fp32_beta_zero_never_reads_c():
    case = Case("R", "N", "N", 70, 90, 50, 1.0, 0.0, seed=3)
    case.C0[:] = np.nan
    C = run_host(case)
    assert np.all(np.isfinite(C))
    case.C0[:] = 0.0
    ok, worst = case.judge(C)
    assert ok


@pytest.mark.parametrize("major", "RC")
def test_alpha_zero_and_k_zero(major):
    case = Case(major, "N", "T", 37, 29, 11, 0.0, 0.5, seed=5, pad=2)
    C = run_host(case)
    row_len = case.N if major == "R" else case.M
    assert np.array_equal(C[:, :row_len], case.C0[:, :row_len] * np.float32(0.5))
    assert case.untouched_padding(C)
    case = Case(major, "N", "N", 37, 29, 0, 1.0, 0.0, seed=6)
    C = run_host(case)
    assert np.all(C[:, :(case.N if major == "R" else case.M)] == 0)
    case = Case(major, "N", "N", 37, 29, 0, 1.0, 1.0, seed=6)
    C = run_host(case)
    assert np.array_equal(C, case.C0)


def test_empty_shapes_are_no_ops():
    A = np.ones(16, np.float32)
    C = np.full(16, 3.0, np.float32)
    ug.sgemm("R", "N", "N", 0, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4)
    ug.sgemm("R", "N", "N", 4, 0, 4, 1.0, A, 4, A, 4, 0.0, C, 4)
    assert np.all(C == 3.0)


def test_argument_errors():
    A = np.ones(64, np.float32)
    with pytest.raises(ug.UgemmError) as e:
        ug.sgemm("R", "N", "N", 8, 8, 8, 1.0, A, 7, A, 8, 0.0, A, 8)
    assert e.value.status == ug.ERR_ARG
    with pytest.raises(ug.UgemmError):
        ug.sgemm("X", "N", "N", 8, 8, 8, 1.0, A, 8, A, 8, 0.0, A, 8)
    with pytest.raises(ug.UgemmError):
        ug.sgemm("R", "N", "N", 8, 8, 8, 1.0, None, 8, A, 8, 0.0, A, 8)


def test_uut_entry_and_packed_shim():
    L = ug.load()
    case = Case("R", "T", "N", 45, 38, 29, 1.5, 0.5, seed=21)
    C = case.C0.copy()
    L.sgemm_hip(b"R", b"T", b"N", case.M, case.N, case.K, case.alpha, case.A.ctypes.data, case.lda,
                case.B.ctypes.data, case.ldb, case.beta, C.ctypes.data, case.ldc)
    assert L.ugemm_hip_last_error() == 0
    assert case.judge(C)[0]
    # Packed shim (sgemm_ocl argument list): ld = K or M, N or K, N.
    for ta, tb in (("N", "N"), ("N", "T"), ("T", "N")):
        case = Case("R", ta, tb, 40, 36, 52, 1.0, 0.0, seed=22)
        C = case.C0.copy()
        L.sgemm_hip_packed(ta.encode(), tb.encode(), case.M, case.N, case.K, case.alpha, case.A.ctypes.data,
                           case.B.ctypes.data, case.beta, C.ctypes.data)
        assert case.judge(C)[0], (ta, tb)


@pytest.mark.parametrize("variant", range(6))
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_fp32_every_variant(variant, ta, tb):
    assert len(ug.variants(ug.MODE_FP32)) == 6
    ug.force_variant(ug.MODE_FP32, variant)
    try:
        for (M, N, K) in ((130, 250, 70), (256, 256, 48), (520, 600, 33)):
            case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=31 + variant)
            C = run_host(case)
            ok, worst = case.judge(C)
            assert ok, f"variant {variant} {M}x{N}x{K}: {worst:.3g}"
    finally:
        ug.force_variant(ug.MODE_FP32, -1)


def test_device_pointer_entry_and_misaligned_bases():
    M, N, K = 300, 260, 200
    case = Case("R", "N", "N", M, N, K, 1.5, 0.5, seed=41)
    s = hip.Stream()
    # One element of offset makes every base 4 mod 16: the vector loads must
    # not be used, results must not change.
    for off in (0, 1):
        dA = hip.DeviceArray(case.A.size + 1, np.float32)
        dB = hip.DeviceArray(case.B.size + 1, np.float32)
        dC = hip.DeviceArray(case.C0.size + 1, np.float32)
        hA = np.zeros(case.A.size + 1, np.float32); hA[off:off + case.A.size] = case.A.ravel()
        hB = np.zeros(case.B.size + 1, np.float32); hB[off:off + case.B.size] = case.B.ravel()
        hC = np.zeros(case.C0.size + 1, np.float32); hC[off:off + case.C0.size] = case.C0.ravel()
        dA.upload(hA); dB.upload(hB); dC.upload(hC)
        ug.sgemm_dev("R", "N", "N", M, N, K, case.alpha, dA.at(off), case.lda, dB.at(off), case.ldb, case.beta,
                     dC.at(off), case.ldc, stream=s)
        s.synchronize()
        C = dC.download()[off:off + case.C0.size].reshape(case.C0.shape)
        ok, worst = case.judge(C)
        assert ok, f"offset {off}: {worst:.3g}"
    s.destroy()


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N")])
def test_fp32_large_sampled(ta, tb):
    # 2048^3 with alpha=1.5, beta=0.5, U[-1,1]: check 64 sampled rows in full.
    M = N = K = 2048
    case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=51)
    C = run_host(case)
    rows = np.sort(orc.rng(52).choice(M, 64, replace=False))
    exact, scale = case.exact(rows=rows)
    ok, worst = orc.check_bound(C[rows, :N], exact, scale, orc.tolerance(K))
    assert ok, f"worst/bound {worst:.3g}"


@pytest.mark.parametrize("mode,beta", [(ug.MODE_FP32, 0.5), (ug.MODE_FP32, 0.0), (ug.MODE_BF16X3, 0.5)])
def test_host_pointer_pipelined_large(mode, beta):
    """Host-pointer calls large enough for the M-block pipeline (copies of
    block b+1 and b-1 overlap the GEMM of block b): every block of C must be
    right, checked on rows sampled from each block."""
    M, N, K = 4100, 4096, 520        # a short last block; 80+ MB moved
    case = Case("R", "N", "N", M, N, K, 1.5, beta, seed=4242)
    C = case.C0.copy()
    ug.sgemm("R", "N", "N", M, N, K, case.alpha, case.A, case.lda, case.B, case.ldb, beta, C, case.ldc, mode=mode)
    rows = np.unique(np.concatenate([np.arange(0, M, 512), np.arange(511, M, 512), [M - 1, M - 4, 2048, 2049]]))
    exact, scale = case.exact(rows=rows)
    from tests.helpers import MODE_EXTRA
    ok, worst = orc.check_bound(C[rows, :N], exact, scale, orc.tolerance(K, extra=MODE_EXTRA.get(mode, 0.0)))
    assert ok, worst


def test_host_pointer_pipelined_dgemm():
    M, N, K = 4096, 2048, 512
    case = Case("R", "N", "N", M, N, K, 1.0, 0.25, seed=4343, dtype=np.float64)
    C = case.C0.copy()
    ug.dgemm("R", "N", "N", M, N, K, 1.0, case.A, case.lda, case.B, case.ldb, 0.25, C, case.ldc)
    rows = np.array([0, 1, 1023, 1024, 2047, 2048, 3071, 4095])
    exact, scale = case.exact(rows=rows)
    ok, worst = orc.check_bound(C[rows, :N], exact, scale, orc.tolerance(K, eps=orc.EPS64))
    assert ok, worst
