"""Parity of the FP64-MFMA DGEMM with the oracle (double-double evaluation of
dgemm_cpu's semantics, ugemm.h:162-286): |C - C_exact| <= 2(K+1) eps64 * scale,
and bitwise determinism across repeated calls."""
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case

pytestmark = pytest.mark.gpu

COMBOS = [(m, a, b) for m in "RC" for a in "NT" for b in "NT"]
SHAPES = [(1, 1, 1), (7, 5, 9), (33, 65, 31), (129, 127, 65), (200, 300, 257)]


def run_host(case):
    C = case.C0.copy()
    ug.dgemm(case.major, case.ta, case.tb, case.M, case.N, case.K, case.alpha, case.A, case.lda, case.B, case.ldb,
             case.beta, C, case.ldc)
    return C


@pytest.mark.parametrize("major,ta,tb", COMBOS)
def test_dgemm_layouts_shapes(major, ta, tb):
    for (M, N, K) in SHAPES:
        for alpha, beta in ((1.0, 0.0), (1.5, 0.5)):
            seed = zlib.crc32(f"d{major}{ta}{tb}{M},{N},{K},{alpha}".encode())
            case = Case(major, ta, tb, M, N, K, alpha, beta, seed=seed, dist="normal", dtype=np.float64)
            C = run_host(case)
            ok, worst = case.judge(C, eps=orc.EPS64)
            assert ok, f"{major}{ta}{tb} {M}x{N}x{K}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("major,ta,tb", COMBOS)
def test_dgemm_matches_naive_reference_loop(major, ta, tb):
    # The reference's own DGEMM oracle is dgemm_cpu (check_dgemm.c:138); its
    # restatement here must agree with the GPU result to 2 K eps64.
    M, N, K = 70, 45, 88
    case = Case(major, ta, tb, M, N, K, 1.5, 0.5, seed=9, dist="normal", dtype=np.float64, pad=3)
    C = run_host(case)
    R = case.C0.copy()
    orc.dgemm_naive(major, ta, tb, M, N, K, 1.5, case.A, case.lda, case.B, case.ldb, 0.5, R, case.ldc)
    _, scale = case.exact()
    ok, worst = orc.check_bound(case.logical(C), case.logical(R).astype(np.float64), scale,
                                orc.tolerance(K, eps=orc.EPS64))
    assert ok, worst
    assert case.untouched_padding(C)


@pytest.mark.parametrize("variant", range(4))
def test_dgemm_every_variant(variant):
    ug.force_variant(ug.MODE_DGEMM, variant)
    try:
        for ta, tb in (("N", "N"), ("T", "T"), ("N", "T"), ("T", "N")):
            case = Case("R", ta, tb, 150, 140, 70, 1.5, 0.5, seed=5 + variant, dist="normal", dtype=np.float64)
            C = run_host(case)
            ok, worst = case.judge(C, eps=orc.EPS64)
            assert ok, f"variant {variant} {ta}{tb}: {worst:.3g}"
    finally:
        ug.force_variant(ug.MODE_DGEMM, -1)


def test_dgemm_integer_exact():
    M, N, K = 100, 90, 300
    g = orc.rng(4)
    A = g.integers(-50, 51, size=(M, K)).astype(np.float64)
    B = g.integers(-50, 51, size=(K, N)).astype(np.float64)
    C = np.zeros((M, N))
    ug.dgemm("R", "N", "N", M, N, K, 1.0, A, K, B, N, 0.0, C, N)
    assert np.array_equal(C, A @ B)


def test_dgemm_deterministic_and_accurate_4096():
    # The DGEMM configuration: 4096^3, normal(0,1) inputs, alpha=1, beta=0.
    n = 4096
    g = orc.rng(0x55474D4D03)
    A = g.standard_normal((n, n))
    B = g.standard_normal((n, n))
    dA, dB = hip.DeviceArray.from_host(A), hip.DeviceArray.from_host(B)
    dC1, dC2 = hip.DeviceArray((n, n), np.float64), hip.DeviceArray((n, n), np.float64)
    ug.dgemm_dev("R", "N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC1, n)
    ug.dgemm_dev("R", "N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC2, n)
    hip.synchronize()
    C1, C2 = dC1.download(), dC2.download()
    assert np.array_equal(C1.view(np.uint64), C2.view(np.uint64)), "DGEMM is not bitwise reproducible"
    rows = np.sort(orc.rng(6).choice(n, 24, replace=False))
    exact, scale = orc.reference_f64("R", "N", "N", n, n, n, 1.0, A, n, B, n, 0.0, None, n, rows=rows)
    ok, worst = orc.check_bound(C1[rows], exact, scale, orc.tolerance(n, eps=orc.EPS64))
    assert ok, worst


def test_dgemm_alpha_zero_and_errors():
    A = np.ones((8, 8))
    C = np.full((8, 8), 2.0)
    ug.dgemm("R", "N", "N", 8, 8, 8, 0.0, A, 8, A, 8, 0.5, C, 8)
    assert np.all(C == 1.0)
    with pytest.raises(ug.UgemmError):
        ug.dgemm("R", "N", "N", 8, 8, 8, 1.0, A, 7, A, 8, 0.0, C, 8)
