"""Split-K: a GEMM whose tile grid leaves CUs idle is split along K, the
partial products of the K chunks are written to a workspace and added in a
fixed order, then alpha, beta and the epilogue are applied.

Checked here for every form that splits (FP32, the split forms, DGEMM):
parity with the oracle to the same bounds as unsplit calls, bitwise equality
of repeated calls (fixed reduction order), beta == 0 never reading C, the
bias/LeakyReLU epilogue after the reduction, the automatic split of a small
grid, and the tuner that times split counts (its version-3 cache file)."""
import os
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case

pytestmark = pytest.mark.gpu

SPLIT_MODES = [ug.MODE_FP32, ug.MODE_BF16X3, ug.MODE_FP16X3, ug.MODE_BF16X1, ug.MODE_BF16X6]
# Few tiles, long K (the shapes split-K is for); K not a multiple of the chunk.
SHAPES = [(200, 130, 3000), (64, 64, 5000), (33, 257, 1100)]


def run_host(case, mode):
    C = case.C0.copy()
    ug.sgemm(case.major, case.ta, case.tb, case.M, case.N, case.K, case.alpha, case.A, case.lda, case.B, case.ldb,
             case.beta, C, case.ldc, mode=mode)
    return C


@pytest.mark.parametrize("mode", SPLIT_MODES)
@pytest.mark.parametrize("splits", [2, 3, 7])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_split_k_parity(mode, splits, ta, tb):
    ug.force_splits(mode, splits)
    try:
        for (M, N, K) in SHAPES:
            seed = zlib.crc32(f"sk{mode}{splits}{ta}{tb}{M},{N},{K}".encode())
            case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=seed)
            C = run_host(case, mode)
            ok, worst = case.judge(C, mode=mode)
            assert ok, f"mode {mode} splits {splits} {ta}{tb} {M}x{N}x{K}: worst/bound {worst:.3g}"
    finally:
        ug.force_splits(mode, 0)


@pytest.mark.parametrize("mode", [ug.MODE_BF16X3, ug.MODE_FP16X3, ug.MODE_BF16X1])
@pytest.mark.parametrize("variant", [4, 6, 7])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_split_k_with_eight_wave_variants(mode, variant, ta, tb):
    """Split K chunks that start at a non-zero K tile, through the 8-wave
    variants: the interleaved stage copy (variant 6, 3 x 3 tiles) and an
    operand split in the core from either FP32 layout (300 or 1100 rows: one
    side of at most two 256-wide tiles; 250: of two 128-wide tiles)."""
    ug.force_variant(mode, variant)
    ug.force_splits(mode, 3)
    try:
        for (M, N, K) in ((600, 700, 3000), (300, 1100, 2500), (1100, 257, 2100), (250, 1100, 2500)):
            case = Case("R", ta, tb, M, N, K, 1.5, 0.5, seed=M + N + K + variant)
            C = run_host(case, mode)
            ok, worst = case.judge(C, mode=mode)
            assert ok, f"variant {variant} {ta}{tb} {M}x{N}x{K}: worst/bound {worst:.3g}"
    finally:
        ug.force_splits(mode, 0)
        ug.force_variant(mode, -1)


@pytest.mark.parametrize("mode", SPLIT_MODES)
def test_split_k_bitwise_reproducible_and_beta_zero(mode):
    M, N, K = 150, 140, 4100
    case = Case("R", "N", "N", M, N, K, 1.0, 0.0, seed=77 + mode)
    dA, dB = hip.DeviceArray.from_host(case.A), hip.DeviceArray.from_host(case.B)
    outs = []
    ug.force_splits(mode, 5)
    try:
        for _ in range(2):
            dC = hip.DeviceArray.from_host(np.full((M, N), np.nan, np.float32))   # beta == 0: never read
            ug.sgemm_dev("R", "N", "N", M, N, K, 1.0, dA, K, dB, N, 0.0, dC, N, mode=mode)
            hip.synchronize()
            outs.append(dC.download())
    finally:
        ug.force_splits(mode, 0)
    assert np.all(np.isfinite(outs[0]))
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    ok, worst = case.judge(outs[0], mode=mode)
    assert ok, worst


@pytest.mark.parametrize("mode", [ug.MODE_FP32, ug.MODE_BF16X3])
def test_split_k_epilogue_after_reduction(mode):
    """Bias and LeakyReLU act on the reduced sum, not on each partial."""
    M, N, K = 96, 200, 3000
    g = orc.rng(31 + mode)
    A, B = orc.uniform(g, (M, K)), orc.uniform(g, (K, N))
    C0, bias = orc.uniform(g, (M, N)), orc.uniform(g, (M,))
    dA, dB, dC, db = (hip.DeviceArray.from_host(x) for x in (A, B, C0, bias))
    ug.force_splits(mode, 4)
    try:
        ug.sgemm_ex_dev("R", "N", "N", M, N, K, 1.25, dA, K, dB, N, 0.5, dC, N, mode=mode, bias=db,
                        bias_dim=ug.BIAS_ROWS, act=ug.ACT_LEAKY_RELU, slope=0.1)
        hip.synchronize()
    finally:
        ug.force_splits(mode, 0)
    got = dC.download().astype(np.float64)
    pre = 1.25 * (A.astype(np.float64) @ B.astype(np.float64)) + 0.5 * C0 + bias[:, None]
    want = np.where(pre >= 0, pre, 0.1 * pre)
    scale = 1.25 * (np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))) + 0.5 * np.abs(C0) + np.abs(bias[:, None])
    extra = {ug.MODE_BF16X3: 2.0 ** -14}.get(mode, 0.0)
    assert np.all(np.abs(got - want) <= orc.tolerance(K, extra=extra) * scale + 1e-30)


def test_split_k_automatic_on_a_small_grid():
    """A 128x128 problem with K = 16384 fills a few tiles: the launcher splits
    it by itself; the result still meets the unsplit bound and repeats bitwise."""
    M, N, K = 128, 128, 16384
    case = Case("R", "N", "N", M, N, K, 1.0, 0.5, seed=5)
    C1, C2 = run_host(case, ug.MODE_FP32), run_host(case, ug.MODE_FP32)
    assert np.array_equal(C1.view(np.uint32), C2.view(np.uint32))
    ok, worst = case.judge(C1)
    assert ok, worst


@pytest.mark.parametrize("splits", [1, 2, 3, 5])
def test_dgemm_split_k_parity_and_determinism(splits):
    ug.force_splits(ug.MODE_DGEMM, splits)
    try:
        for ta, tb in (("N", "N"), ("T", "N"), ("N", "T")):
            case = Case("R", ta, tb, 100, 90, 4000, 1.5, 0.5, seed=splits * 7 + len(ta + tb), dist="normal",
                        dtype=np.float64)
            C1, C2 = case.C0.copy(), case.C0.copy()
            for C in (C1, C2):
                ug.dgemm("R", ta, tb, case.M, case.N, case.K, 1.5, case.A, case.lda, case.B, case.ldb, 0.5, C,
                         case.ldc)
            assert np.array_equal(C1.view(np.uint64), C2.view(np.uint64))
            ok, worst = case.judge(C1, eps=orc.EPS64)
            assert ok, f"splits {splits} {ta}{tb}: {worst:.3g}"
    finally:
        ug.force_splits(ug.MODE_DGEMM, 0)


def test_force_splits_argument_errors():
    L = ug.load()
    assert L.ugemm_hip_force_splits(ug.MODE_FP32, -1) == ug.ERR_ARG
    assert L.ugemm_hip_force_splits(ug.MODE_FP32, 65) == ug.ERR_ARG
    assert L.ugemm_hip_force_splits(99, 2) == ug.ERR_ARG
    assert L.ugemm_hip_force_splits(ug.MODE_FP32, 0) == ug.OK


def test_tuner_times_split_counts_and_persists_them(tmp_path):
    """ugemm_hip_tune_sgemm on a few-tile, long-K shape: the chosen variant is
    recorded with a split count; the cache file is version 3 and loads back."""
    L = ug.load()
    M, N, K = 256, 192, 8192
    v = L.ugemm_hip_tune_sgemm(b"N", b"N", M, N, K, ug.MODE_FP32)
    assert 0 <= v < len(ug.variants(ug.MODE_FP32))
    assert L.ugemm_hip_tune_lookup(b"N", b"N", M, N, K, ug.MODE_FP32) == v
    path = os.path.join(str(tmp_path), "tune.txt")
    assert L.ugemm_hip_tune_save(path.encode()) == ug.OK
    lines = open(path).read().split("\n")
    assert lines[0].startswith("ugemm-hip-tune 3 ")
    rows = [l.split() for l in lines[1:] if l.strip()]
    mine = [r for r in rows if int(r[0]) == ug.MODE_FP32 and int(r[6]) == v]
    assert mine and all(len(r) == 9 and 0 <= int(r[8]) <= 64 for r in rows)
    assert L.ugemm_hip_tune_load(path.encode()) == len(rows)
    # The tuned entry is used by the next call on that bucket.
    case = Case("R", "N", "N", M, N, K, 1.0, 0.0, seed=3)
    ok, worst = case.judge(run_host(case, ug.MODE_FP32))
    assert ok, worst


def test_tune_cache_environment_variable_loaded_at_init(tmp_path):
    """UGEMM_HIP_TUNE_CACHE: a fresh process that only initialises the library
    (as a uut caller of sgemm_hip would) sees the saved table's entry; a
    missing file is ignored."""
    import subprocess
    import sys
    L = ug.load()
    M, N, K = 300, 200, 500
    v = L.ugemm_hip_tune_sgemm(b"N", b"N", M, N, K, ug.MODE_FP32)
    assert v >= 0
    path = os.path.join(str(tmp_path), "tune.txt")
    assert L.ugemm_hip_tune_save(path.encode()) == ug.OK
    # Rewrite that bucket's variant to another valid one, so that the loaded
    # entry cannot be mistaken for the built-in choice.
    want = (v + 1) % len(ug.variants(ug.MODE_FP32))
    lines = open(path).read().split("\n")
    lg = lambda x: int(np.ceil(np.log2(x)))
    for i, l in enumerate(lines[1:], 1):
        f = l.split()
        if f and [int(t) for t in f[:6]] == [ug.MODE_FP32, 0, 0, lg(M), lg(N), lg(K)]:
            f[6] = str(want)
            lines[i] = " ".join(f)
    open(path, "w").write("\n".join(lines))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import ugemm_amd as ug; L = ug.load(); ug.init(); "
            f"print(L.ugemm_hip_tune_lookup(b'N', b'N', {M}, {N}, {K}, 0))")
    for cache, expect in ((path, want), (os.path.join(str(tmp_path), "missing.txt"), -1)):
        r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=120,
                           env={**os.environ, "UGEMM_HIP_TUNE_CACHE": cache})
        assert r.returncode == 0, r.stderr
        assert int(r.stdout.split()[-1]) == expect, (cache, r.stdout)
