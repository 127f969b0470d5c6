"""Parity at the benchmark configurations' own sizes (sampled rows or blocks
of C against the oracle, each form with its own bound):

  * config 3: SGEMM 65536 x 256 x 4096 and 256 x 65536 x 4096, normal(0, 1)
    inputs, in the FP32, FP32-MFMA, BF16x3, FP16X3 and BF16x6 forms;
  * config 4: the 512-problem im2col draw on independent operands (one
    device span per operand, 7.5 GB resident), 2 sampled rows of every
    problem, in FP32, FP32-MFMA, BF16x3 and FP16X3 grouped plans;
  * config 5: the wide-range precision ladder (entries sign * 2^e * (1+u),
    e in [-12, 12]) at 4096^3, every form against its oracle: the FP64
    product for the FP32-class forms, the exact emulation of the split for
    BF16X1/X3/X6 and FP16X3, the block-quantisation emulation for FP8;
  * the split forms' NT and TN layouts at 2048^3 (FP16X3 and FP8 against
    their emulations, FP32-MFMA against FP64).
Shapes follow the conv -> GEMM mapping of sgemm_ocl1.h:271-300 (config 4) and
the semantics of ugemm.h:287-411."""
import os
import sys
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import MODE_EXTRA

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (config 4's shape draw and operand generator)

pytestmark = pytest.mark.gpu

FP8_EXTRA = 2.0 ** -12   # the FP8 MFMA's internal summation (tests/test_gpu_fp8.py)


def device_gemm(A, B, M, N, K, ta, tb, mode, alpha=1.0, beta=0.0, C0=None):
    """One ugemm_hip_sgemm_dev call on fresh device copies; returns C."""
    lda = A.shape[1]
    ldb = B.shape[1]
    dA, dB = hip.DeviceArray.from_host(A), hip.DeviceArray.from_host(B)
    dC = hip.DeviceArray.from_host(C0) if C0 is not None else hip.DeviceArray((M, N), np.float32)
    try:
        ug.sgemm_dev("R", ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, N, mode=mode)
        return dC.download()
    finally:
        for d in (dA, dB, dC):
            d.free()


# ---------------------------------------------------------------------------
# Config 3: tall-skinny and short-wide.

C3_FORMS = [ug.MODE_FP32, ug.MODE_FP32_MFMA, ug.MODE_BF16X3, ug.MODE_FP16X3, ug.MODE_BF16X6]


@pytest.fixture(scope="module")
def config3_operands():
    g = np.random.Generator(np.random.PCG64(0x55474D4D03))
    big = g.standard_normal((65536, 4096), dtype=np.float32)
    small = g.standard_normal((4096, 256), dtype=np.float32)
    return big, small


@pytest.mark.parametrize("mode", C3_FORMS)
def test_config3_tall_skinny(config3_operands, mode):
    """65536 x 256 x 4096 NN, alpha 1, beta 0: 32 sampled rows vs FP64."""
    big, small = config3_operands
    M, N, K = 65536, 256, 4096
    print(f"config3 tall-skinny mode {mode}", flush=True)
    C = device_gemm(big, small, M, N, K, "N", "N", mode)
    rows = np.sort(orc.rng(31).choice(M, 32, replace=False))
    Ar = big[rows].astype(np.float64)
    exact = Ar @ small.astype(np.float64)
    scale = np.abs(Ar) @ np.abs(small.astype(np.float64))
    ok, worst = orc.check_bound(C[rows], exact, scale, orc.tolerance(K, extra=MODE_EXTRA.get(mode, 0.0)))
    assert ok, f"mode {mode}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("mode", C3_FORMS)
def test_config3_short_wide(config3_operands, mode):
    """256 x 65536 x 4096 NN (A = small^T, B = big^T), alpha 1, beta 0: a
    32-row x 2048-column sample of C vs FP64."""
    big, small = config3_operands
    M, N, K = 256, 65536, 4096
    A = np.ascontiguousarray(small.T)
    B = np.ascontiguousarray(big.T)
    print(f"config3 short-wide mode {mode}", flush=True)
    C = device_gemm(A, B, M, N, K, "N", "N", mode)
    g = orc.rng(32)
    rows = np.sort(g.choice(M, 32, replace=False))
    cols = np.sort(g.choice(N, 2048, replace=False))
    Ar, Bc = A[rows].astype(np.float64), B[:, cols].astype(np.float64)
    exact, scale = Ar @ Bc, np.abs(Ar) @ np.abs(Bc)
    ok, worst = orc.check_bound(C[np.ix_(rows, cols)], exact, scale,
                                orc.tolerance(K, extra=MODE_EXTRA.get(mode, 0.0)))
    assert ok, f"mode {mode}: worst/bound {worst:.3g}"


# ---------------------------------------------------------------------------
# Config 4: the 512-problem batch on independent operands.

@pytest.fixture(scope="module")
def config4_batch():
    _, Ms, HW, Ks = bench.config4_shapes()
    Ns = HW * HW
    dA, dB, dC, a_off, b_off, c_off, checks, resident = bench.config4_operands(Ms, Ns, Ks)
    probs = [ug.Problem(b"N", b"N", int(Ms[i]), int(Ns[i]), int(Ks[i]), 1.0, 0.0, dA.at(a_off[i]), int(Ks[i]),
                        dB.at(b_off[i]), int(Ns[i]), dC.at(c_off[i]), int(Ns[i])) for i in range(len(Ms))]
    yield Ms, Ns, Ks, probs, dC, c_off, checks, resident
    for d in (dA, dB, dC):
        d.free()


@pytest.mark.parametrize("mode", [ug.MODE_FP32, ug.MODE_FP32_MFMA, ug.MODE_BF16X3, ug.MODE_FP16X3])
def test_config4_every_problem(config4_batch, mode):
    Ms, Ns, Ks, probs, dC, c_off, checks, resident = config4_batch
    assert len(probs) == 512 and resident > 7e9
    assert Ms.min() >= 32 and Ms.max() <= 512 and set(np.unique(Ns)) <= {49, 196, 784, 3136}
    assert Ks.min() >= 144 and Ks.max() <= 4608 and np.all(Ks % 144 == 0)
    print(f"config4 512 problems mode {mode}", flush=True)
    hip.check(hip.lib().hipMemset(dC.ptr, 0xFF, dC.nbytes), "poison C")   # NaN: every C must be written
    plan = ug.GroupedPlan(probs, mode=mode, capture=True)
    try:
        plan.run()
        plan.run()          # a replay overwrites with the same values
        hip.synchronize()
    finally:
        plan.destroy()
    worst = bench.config4_check(dC, c_off, Ms, Ns, Ks, checks, MODE_EXTRA.get(mode, 0.0))
    assert worst <= 1.0, f"mode {mode}: worst/bound {worst:.3g}"


# ---------------------------------------------------------------------------
# Config 5: the wide-range ladder at 4096^3.

@pytest.fixture(scope="module")
def config5_operands():
    g = np.random.Generator(np.random.PCG64(0x55474D4D05))
    n = 4096
    return n, bench.wide(g, (n, n)), bench.wide(g, (n, n))


LADDER = [ug.MODE_FP32, ug.MODE_FP32_MFMA, ug.MODE_BF16X3, ug.MODE_BF16X1, ug.MODE_BF16X6, ug.MODE_FP16X3,
          ug.MODE_FP8_BLOCK]


@pytest.mark.parametrize("mode", LADDER)
def test_config5_wide_range_ladder(config5_operands, mode):
    n, A, B = config5_operands
    print(f"config5 4096^3 wide range mode {mode}", flush=True)
    C = device_gemm(A, B, n, n, n, "N", "N", mode)
    rows = np.sort(orc.rng(55).choice(n, 16, replace=False))
    if mode == ug.MODE_FP8_BLOCK:
        # Block scales need every row of a sampled row's 128-row block.
        blocks = np.unique(rows // 128)
        Aq = np.concatenate([orc.fp8_block_dequant(A[b * 128:(b + 1) * 128]) for b in blocks])
        sel = np.concatenate([np.nonzero(np.isin(np.arange(b * 128, (b + 1) * 128), rows))[0] + 128 * i
                              for i, b in enumerate(blocks)])
        Aq = Aq[sel]
        Bq = orc.fp8_block_dequant(B.T).T
        ref, scale = Aq @ Bq, np.abs(Aq) @ np.abs(Bq)
        tol = orc.tolerance(n, extra=FP8_EXTRA)
    elif mode in (ug.MODE_BF16X1, ug.MODE_BF16X3, ug.MODE_BF16X6):
        NP = {ug.MODE_BF16X1: 1, ug.MODE_BF16X3: 2, ug.MODE_BF16X6: 3}[mode]
        ref = orc.bf16_product(A[rows], B, NP)
        scale = np.abs(A[rows].astype(np.float64)) @ np.abs(B.astype(np.float64))
        tol = orc.tolerance(n)
    elif mode == ug.MODE_FP16X3:
        ref = orc.fp16x3_product(A[rows], B)
        scale = np.abs(A[rows].astype(np.float64)) @ np.abs(B.astype(np.float64))
        tol = orc.tolerance(n)
    else:
        Ar = A[rows].astype(np.float64)
        ref, scale = Ar @ B.astype(np.float64), np.abs(Ar) @ np.abs(B.astype(np.float64))
        tol = orc.tolerance(n)
    ok, worst = orc.check_bound(C[rows], ref, scale, tol)
    assert ok, f"mode {mode}: worst/bound {worst:.3g}"


# ---------------------------------------------------------------------------
# Split forms in NT and TN at 2048^3 (BF16X1/X3/X6 are in test_gpu_bf16_emulation.py).

@pytest.mark.parametrize("mode", [ug.MODE_FP16X3, ug.MODE_FP8_BLOCK, ug.MODE_FP32_MFMA])
@pytest.mark.parametrize("ta,tb", [("N", "T"), ("T", "N"), ("T", "T")])
def test_split_forms_transposed_2048(mode, ta, tb):
    n = 2048
    g = orc.rng(zlib.crc32(f"t2048{mode}{ta}{tb}".encode()))
    A, B, C0 = orc.uniform(g, (n, n)), orc.uniform(g, (n, n)), orc.uniform(g, (n, n))
    alpha, beta = 1.5, 0.5
    print(f"2048^3 {ta}{tb} mode {mode}", flush=True)
    C = device_gemm(A, B, n, n, n, ta, tb, mode, alpha, beta, C0)
    rows = np.sort(g.choice(n, 24, replace=False))
    Aop = orc.op_matrix("R", ta, A, n, n)
    Bop = orc.op_matrix("R", tb, B, n, n)
    C0r = C0[rows].astype(np.float64)
    if mode == ug.MODE_FP16X3:
        P = orc.fp16x3_product(Aop[rows], Bop)
        Pscale = np.abs(Aop[rows].astype(np.float64)) @ np.abs(Bop.astype(np.float64))
        tol = orc.tolerance(n)
    elif mode == ug.MODE_FP8_BLOCK:
        Aq = orc.fp8_block_dequant(Aop)[rows]
        Bq = orc.fp8_block_dequant(Bop.T).T
        P, Pscale = Aq @ Bq, np.abs(Aq) @ np.abs(Bq)
        tol = orc.tolerance(n, extra=FP8_EXTRA)
    else:
        Ar = Aop[rows].astype(np.float64)
        P, Pscale = Ar @ Bop.astype(np.float64), np.abs(Ar) @ np.abs(Bop.astype(np.float64))
        tol = orc.tolerance(n)
    ref = alpha * P + beta * C0r
    scale = abs(alpha) * Pscale + abs(beta) * np.abs(C0r)
    ok, worst = orc.check_bound(C[rows], ref, scale, tol)
    assert ok, f"mode {mode} {ta}{tb}: worst/bound {worst:.3g}"
