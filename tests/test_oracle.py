"""Pins the CPU oracle (oracle/oracle.c, oracle/oracle.py) -- CPU only.

Checks against
  * the known-answer product of sgemm_test.c:186-200;
  * the reference's own outputs (tests/golden/ref_gemm.npz, written by
    tests/golden/make_golden.py from the reference compiled from its sources);
  * exact small-integer products and numpy's FP64 matmul (an independent
    library routine);
  * planted mistakes (transposed operand, dropped term, wrong sign, wrong
    beta): the acceptance check must reject each of them;
  * the e4m3 definition (known encodings, ties, subnormals) for the FP8 emulation.
"""
import ast
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
COMBOS = [(m, a, b) for m in "RC" for a in "NT" for b in "NT"]


def make(major, ta, tb, M, N, K, dtype=np.float32, pad=0, seed=1, gen=orc.uniform):
    row = major == "R"
    la = (M if ta == "T" else K) if row else (K if ta == "T" else M)
    lb = (K if tb == "T" else N) if row else (N if tb == "T" else K)
    lc = N if row else M
    lda, ldb, ldc = la + pad, lb + pad, lc + pad
    g = orc.rng(seed)
    A = gen(g, orc.stored_shape(major, ta, M, K, lda), dtype=dtype)
    B = gen(g, orc.stored_shape(major, tb, K, N, ldb), dtype=dtype)
    C = gen(g, (M, ldc) if row else (N, ldc), dtype=dtype)
    return A, lda, B, ldb, C, ldc


# ---------------------------------------------------------------------------
def test_known_answer_all_oracle_paths():
    A = np.arange(1, 7, dtype=np.float32)
    B = np.arange(1, 7, dtype=np.float32)
    want = [9, 12, 15, 19, 26, 33, 29, 40, 51]
    for fn in (orc.sgemm_naive, orc.sgemm_packed):
        C = np.zeros(9, np.float32)
        fn("R", "N", "N", 3, 3, 2, 1.0, A, 2, B, 3, 0.0, C, 3)
        assert C.tolist() == want
    out, scale = orc.reference_f64("R", "N", "N", 3, 3, 2, 1.0, A, 2, B, 3, 0.0, None, 3)
    assert out.ravel().tolist() == want


@pytest.mark.parametrize("major,ta,tb", COMBOS)
def test_naive_and_packed_match_numpy_fp64(major, ta, tb):
    M, N, K = 23, 17, 41
    for dtype in (np.float32, np.float64):
        A, lda, B, ldb, C0, ldc = make(major, ta, tb, M, N, K, dtype=dtype, pad=3, seed=7)
        opA = orc.op_matrix(major, ta, A, M, K).astype(np.float64)
        opB = orc.op_matrix(major, tb, B, K, N).astype(np.float64)
        want = 1.5 * (opA @ opB) + 0.5 * orc.c_matrix(major, C0, M, N).astype(np.float64)
        scale = 1.5 * (np.abs(opA) @ np.abs(opB)) + 0.5 * np.abs(orc.c_matrix(major, C0, M, N))
        eps = orc.EPS32 if dtype == np.float32 else orc.EPS64
        for fn in ((orc.sgemm_naive, orc.sgemm_packed) if dtype == np.float32 else (orc.dgemm_naive, orc.dgemm_packed)):
            C = C0.copy()
            fn(major, ta, tb, M, N, K, 1.5, A, lda, B, ldb, 0.5, C, ldc)
            ok, worst = orc.check_bound(orc.c_matrix(major, C, M, N), want, scale, orc.tolerance(K, eps=eps))
            assert ok, (fn.__name__, worst)
            row_len = N if major == "R" else M
            assert np.array_equal(C[:, row_len:], C0[:, row_len:]), "padding written"
        exact, sc = orc.reference_f64(major, ta, tb, M, N, K, 1.5, A, lda, B, ldb, 0.5, C0, ldc)
        assert np.allclose(exact, want, rtol=1e-12, atol=1e-12 * np.abs(scale).max())
        assert np.allclose(sc, scale, rtol=1e-12)


def test_reference_f64_is_exact_on_integers_and_row_subsets():
    M, N, K = 40, 30, 500
    g = orc.rng(2)
    A = g.integers(-1000, 1001, size=(M, K)).astype(np.float32)
    B = g.integers(-1000, 1001, size=(K, N)).astype(np.float32)
    exact = A.astype(np.int64) @ B.astype(np.int64)
    out, _ = orc.reference_f64("R", "N", "N", M, N, K, 1.0, A, K, B, N, 0.0, None, N)
    assert np.array_equal(out, exact.astype(np.float64))
    rows = np.array([3, 17, 39])
    sub, _ = orc.reference_f64("R", "N", "N", M, N, K, 1.0, A, K, B, N, 0.0, None, N, rows=rows)
    assert np.array_equal(sub, out[rows])
    # double-double: products of 2^30-range integers summed exactly
    Ad = g.integers(-(1 << 30), 1 << 30, size=(8, 64)).astype(np.float64)
    Bd = g.integers(-(1 << 30), 1 << 30, size=(64, 8)).astype(np.float64)
    outd, _ = orc.reference_f64("R", "N", "N", 8, 8, 64, 1.0, Ad, 64, Bd, 8, 0.0, None, 8)
    exact_d = [[sum(int(Ad[i, k]) * int(Bd[k, j]) for k in range(64)) for j in range(8)] for i in range(8)]
    assert np.array_equal(outd, np.array(exact_d, dtype=np.float64))


def test_planted_mistakes_are_rejected():
    # The acceptance check must fail for plausible kernel mistakes.
    M, N, K = 48, 40, 64
    A, lda, B, ldb, C0, ldc = make("R", "N", "N", M, N, K, seed=3)
    exact, scale = orc.reference_f64("R", "N", "N", M, N, K, 1.5, A, lda, B, ldb, 0.5, C0, ldc)
    tol = orc.tolerance(K)
    good = C0.copy()
    orc.sgemm_naive("R", "N", "N", M, N, K, 1.5, A, lda, B, ldb, 0.5, good, ldc)
    assert orc.check_bound(good, exact, scale, tol)[0]
    A64, B64, C64 = A.astype(np.float64), B.astype(np.float64), C0.astype(np.float64)
    mistakes = {
        "B transposed": 1.5 * (A64[:, :N] @ B64[:N, :N].T[:, :N]) if K >= N else None,
        "dropped last k": 1.5 * (A64[:, :-1] @ B64[:-1]) + 0.5 * C64,
        "wrong sign of beta term": 1.5 * (A64 @ B64) - 0.5 * C64,
        "beta ignored": 1.5 * (A64 @ B64),
        "alpha applied twice": 2.25 * (A64 @ B64) + 0.5 * C64,
        "k index shifted": 1.5 * (A64[:, 1:] @ B64[:-1]) + 0.5 * C64,
    }
    for name, bad in mistakes.items():
        if bad is None:
            continue
        assert not orc.check_bound(bad, exact, scale, tol)[0], name


def test_cmp_results_statistic():
    ref = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    res = ref.copy()
    res[1, 0] += 1e-3
    std_err, std_ref, max_err, (i, j), verdict = orc.cmp_results(2, 2, ref, res, 2)
    assert (i, j) == (1, 0) and abs(max_err - 1e-3) < 1e-6
    assert abs(std_ref - np.sqrt(1.25)) < 1e-12
    assert verdict == "FAIL !!!"


def test_generators():
    g = orc.rng(5)
    u = orc.uniform(g, 10000)
    k = (u.astype(np.float64) + 1.0) * 2.0 ** 23
    assert np.all(k == np.round(k)) and u.min() >= -1.0 and u.max() < 1.0
    w = orc.wide_range(orc.rng(6), 10000)
    e = np.floor(np.log2(np.abs(w.astype(np.float64))))
    assert e.min() == -12 and e.max() == 12


# ---------------------------------------------------------------------------
# Golden vectors from the compiled reference.

def load_golden():
    meta = json.load(open(os.path.join(GOLDEN, "ref_gemm.json")))
    data = np.load(os.path.join(GOLDEN, "ref_gemm.npz"))
    return meta["cases"], data


def golden_inputs(case, data):
    if case["inputs"] is not None:
        p = case["inputs"]
        return data[p + "_A"], data[p + "_B"], data[p + "_C0"]
    dtype = np.dtype(case["dtype"])
    g = orc.rng(case["seed"])
    sa = orc.stored_shape(case["major"], case["ta"], case["M"], case["K"], case["lda"])
    sb = orc.stored_shape(case["major"], case["tb"], case["K"], case["N"], case["ldb"])
    sc = (case["M"], case["ldc"]) if case["major"] == "R" else (case["N"], case["ldc"])
    A, B, C0 = orc.uniform(g, sa, dtype=dtype), orc.uniform(g, sb, dtype=dtype), orc.uniform(g, sc, dtype=dtype)
    if hashlib.sha256(A.tobytes() + B.tobytes() + C0.tobytes()).hexdigest() != case["inputs_sha256"]:
        pytest.skip("regenerated inputs differ from the fixture's (numpy stream changed)")
    return A, B, C0


def reference_divergence(case):
    """Why the reference's result deliberately differs from BLAS semantics
    (which the oracle and the GPU path follow), or None."""
    fn, major = case["fn"], case["major"]
    blocked = ("sgemm_c", "_dgemm_c", "ref_sgemm_sse", "dgemm_avx", "sgemm_avx")
    scales_only = (case["alpha"] == 0.0 or case["K"] == 0) and case["beta"] != 1.0
    if scales_only:
        if fn in blocked and major == "R" and case["M"] != case["N"]:
            return "alpha == 0 indexes C column-major (gemm_cpu.h:308-323)"
        return None
    if fn == "sgemm_avx" and (case["ta"] != "N" or case["tb"] != "N"):
        return "sgemm_avx implements R/N/N only (sgemm_avx256.h:439-453)"
    return None


def test_golden_fixture_integrity():
    meta = json.load(open(os.path.join(GOLDEN, "ref_gemm.json")))
    h = hashlib.sha256(open(os.path.join(GOLDEN, "ref_gemm.npz"), "rb").read()).hexdigest()
    assert h == meta["sha256"]
    fns = {c["fn"] for c in meta["cases"]}
    assert fns == {"sgemm_cpu", "sgemm_c", "sgemm_avx", "ref_sgemm_sse", "dgemm_cpu", "_dgemm_c", "dgemm_avx"}


def test_golden_known_answer():
    cases, data = load_golden()
    kat = [i for i, c in enumerate(cases) if c["M"] == 3 and c["N"] == 3 and c["K"] == 2 and c["inputs"] is not None
           and np.array_equal(data[c["inputs"] + "_A"].ravel(), np.arange(1, 7))]
    assert len(kat) >= 4
    for i in kat:
        assert data[f"C{i}"].ravel().tolist() == [9, 12, 15, 19, 26, 33, 29, 40, 51]


def test_oracle_matches_reference_golden_vectors():
    cases, data = load_golden()
    checked = diverged = 0
    for i, c in enumerate(cases):
        A, B, C0 = golden_inputs(c, data)
        ref_out = data[f"C{i}"]
        why = reference_divergence(c)
        eps = orc.EPS32 if c["dtype"] == "float32" else orc.EPS64
        args = (c["major"], c["ta"], c["tb"], c["M"], c["N"], c["K"])
        exact, scale = orc.reference_f64(*args, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C0, c["ldc"])
        got_ref = orc.c_matrix(c["major"], ref_out, c["M"], c["N"])
        ok, worst = orc.check_bound(got_ref, exact, scale, orc.tolerance(c["K"], eps=eps))
        if why is None:
            assert ok, (c, worst)
            assert not c["writes_outside_C"], c
            checked += 1
        else:
            if "R/N/N only" in why:
                assert np.array_equal(ref_out, C0), c       # C left untouched
            diverged += 1
        # The naive restatement follows sgemm_cpu / dgemm_cpu exactly in
        # semantics (same index formulas, C read even when beta == 0).
        if c["fn"] in ("sgemm_cpu", "dgemm_cpu"):
            mine = C0.copy()
            (orc.sgemm_naive if c["dtype"] == "float32" else orc.dgemm_naive)(
                *args, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], mine, c["ldc"])
            ok2, w2 = orc.check_bound(orc.c_matrix(c["major"], mine, c["M"], c["N"]), got_ref.astype(np.float64),
                                      scale, orc.tolerance(c["K"], eps=eps))
            assert ok2, (c, w2)
    assert checked > 1000 and diverged > 0


def test_reference_alpha_zero_bug_is_real():
    # The documented divergence: the reference's blocked paths write outside
    # the row-major C for alpha == 0 and M != N (captured with a canary).
    cases, _ = load_golden()
    bad = [c for c in cases if c["writes_outside_C"]]
    assert bad and all(c["alpha"] == 0.0 and c["major"] == "R" and c["M"] < c["N"] for c in bad)


# ---------------------------------------------------------------------------
# e4m3 (FP8) emulation.

def test_e4m3_table_and_rounding():
    v = orc.e4m3_table()
    assert len(v) == 127 and v[-1] == 448.0 and v[1] == 2.0 ** -9 and v[8] == 2.0 ** -6
    q = orc.quantize_e4m3
    assert q(448.0) == 448.0 and q(460.0) == 448.0
    assert q(17.0) == 16.0                        # tie between 16 and 18 -> even mantissa
    assert q(19.0) == 20.0                        # tie between 18 and 20 -> 20 (mantissa 2 vs 1)
    assert q(1.0625) == 1.0 and q(1.1875) == 1.25
    assert q(2.0 ** -10) == 0.0                   # half the smallest subnormal: tie -> 0
    assert q(3 * 2.0 ** -10) == 2.0 ** -8         # tie between 1 and 2 subnormal steps -> 2
    assert q(-0.3) == -0.3125
    assert orc.block_exponent(448.0) == 0 and orc.block_exponent(449.0) == 1
    assert orc.block_exponent(1.0) == -8 and orc.block_exponent(0.0) == 0


def test_fp8_block_dequant_uses_one_scale_per_block():
    X = np.zeros((130, 256))
    X[0, 0] = 448.0 * 4          # block (0, 0): scale 4
    X[0, 1] = 5.0                # quantised with the block scale: 5/4 -> 1.25 -> 5.0
    X[129, 200] = 1.0            # block (1, 1): scale 2^-8
    X[129, 201] = 0.3
    D = orc.fp8_block_dequant(X)
    assert D[0, 0] == 1792.0 and D[0, 1] == 5.0
    assert D[129, 200] == 1.0 and D[129, 201] == orc.quantize_e4m3(0.3 * 256) / 256


# ---------------------------------------------------------------------------
# FP16X3 emulation.

def _half_rne(x):
    """IEEE binary16 rounding of a float by Python's struct codec (an
    implementation independent of numpy's casts)."""
    import struct
    return struct.unpack("<e", struct.pack("<e", float(x)))[0]


def test_fp16_split_matches_struct_half_and_reconstructs():
    g = orc.rng(91)
    X = orc.wide_range(g, (6, 64), emin=-20, emax=20)
    X[0, :3] = [1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, -(2.0 ** 14 + 2.0 ** 3)]   # ties to even
    hi, lo, e = orc.fp16_split_rows(X)
    for r in range(X.shape[0]):
        xs = np.ldexp(X[r].astype(np.float64), -e[r])
        assert np.max(np.abs(xs)) < 2.0 ** 15 and np.max(np.abs(xs)) >= 2.0 ** 14
        for k in range(0, X.shape[1], 7):
            h = _half_rne(xs[k])
            assert hi[r, k] == h
            assert lo[r, k] == _half_rne(np.float32(xs[k] - h))
        # hi + lo within 2^-22 relative (or the fp16 subnormal step 2^-25 absolute)
        err = np.abs(xs - hi[r] - lo[r])
        assert np.all(err <= np.maximum(2.0 ** -22 * np.abs(xs), 2.0 ** -25))


def test_fp16_row_exponent_special_cases():
    assert orc.fp16_row_exponent(0.0) == -15
    assert orc.fp16_row_exponent(np.inf) == -15 and orc.fp16_row_exponent(np.nan) == -15
    assert orc.fp16_row_exponent(1.0) == -14            # 1 * 2^14 in [2^14, 2^15)
    assert orc.fp16_row_exponent(2.0 ** 15 - 1) == 0
    assert orc.fp16_row_exponent(1e-30) == int(np.frexp(1e-30)[1]) - 15


def test_fp16x3_product_exact_on_small_integers_and_catches_dropped_terms():
    g = orc.rng(92)
    A = g.integers(-1024, 1025, size=(9, 40)).astype(np.float32)
    B = g.integers(-1024, 1025, size=(40, 11)).astype(np.float32)
    assert np.array_equal(orc.fp16x3_product(A, B), A.astype(np.float64) @ B.astype(np.float64))
    # On general data the emulation meets the 2^-19 per-product bound, and the
    # hi*hi term alone does not (the bound would catch a dropped product).
    A = orc.uniform(g, (32, 512)); B = orc.uniform(g, (512, 24))
    exact = A.astype(np.float64) @ B.astype(np.float64)
    scale = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    assert np.all(np.abs(orc.fp16x3_product(A, B) - exact) <= 2.0 ** -19 * scale)
    ah, _, ea = orc.fp16_split_rows(A)
    bh, _, eb = orc.fp16_split_rows(B.T)
    hi_only = np.ldexp(ah @ bh.T, ea[:, None] + eb[None, :])
    assert not np.all(np.abs(hi_only - exact) <= 2.0 ** -19 * scale)


# ---------------------------------------------------------------------------
# BF16 split emulation (the BF16X1 / BF16X3 / BF16X6 forms).

def _bf16_nearest(x):
    """Independent restatement of round-to-nearest-even onto bf16 for one
    finite float32: the two bf16 neighbours of x are its float32 bit pattern
    truncated to the top 16 bits and the next bf16 away from zero; the nearer
    wins, a tie goes to the one whose last kept bit (bit 16) is 0."""
    u = int(np.float32(x).view(np.uint32))
    lo = u & 0xFFFF0000
    hi = lo + 0x10000
    v_lo = float(np.uint32(lo).view(np.float32))
    v_hi = float(np.uint32(hi).view(np.float32)) if ((hi >> 23) & 0xFF) != 0xFF else float("inf") * np.sign(x)
    d_lo, d_hi = abs(float(x) - v_lo), abs(v_hi - float(x))
    if d_lo < d_hi or (d_lo == d_hi and (lo >> 16) % 2 == 0):
        return v_lo
    return v_hi


def _trunc_bf16(x):
    x = np.asarray(x, np.float32)
    return (x.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)


def _trunc_split_product(Aop, Bop, NP):
    """A planted mistake: the BF16 split with truncation instead of rounding."""
    def split(X):
        X = np.asarray(X, np.float32)
        h = _trunc_bf16(X)
        out = [h.astype(np.float64)]
        r = (X - h).astype(np.float32)
        if NP > 1:
            l1 = _trunc_bf16(r)
            out.append(l1.astype(np.float64))
            if NP > 2:
                out.append(_trunc_bf16((r - l1).astype(np.float32)).astype(np.float64))
        return out
    ap, bp = split(Aop), split(np.asarray(Bop, np.float32).T)
    return sum(ap[i] @ bp[j].T for (i, j) in orc.bf16_products(NP))


def test_bf16_round_matches_nearest_value_search():
    g = orc.rng(95)
    bits = g.integers(0, 0x7F7F0000, size=4000, dtype=np.int64).astype(np.uint32)   # every exponent, subnormals too
    xs = list(bits.view(np.float32)) + list(-bits[:500].view(np.float32))
    # exact ties (a 1 in bit 15 and zeros below) on both parities of bit 16, and values just off a tie
    for m in (0x3F800000, 0x3F810000, 0x40490000, 0x00010000, 0x7F000000):
        xs += [np.uint32(m | 0x8000).view(np.float32), np.uint32(m | 0x8001).view(np.float32),
               np.uint32(m | 0x7FFF).view(np.float32)]
    xs = np.array(xs, np.float32)
    got = orc.bf16_round(xs)
    want = np.array([_bf16_nearest(x) for x in xs], np.float64)
    assert np.array_equal(got.astype(np.float64), want)
    assert np.all((got.view(np.uint32) & 0xFFFF) == 0)
    assert orc.bf16_round(np.float32(1.0 + 2.0 ** -8)) == 1.0                 # tie to even (down)
    assert orc.bf16_round(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6  # tie to even (up)
    sp = orc.bf16_round(np.array([np.inf, -np.inf, np.nan], np.float32))
    assert sp[0] == np.inf and sp[1] == -np.inf and np.isnan(sp[2])


def test_bf16_split_reconstruction_bounds_are_tight():
    """|x - h0| <= 2^-8|x| and |x - h0 - h1| <= 2^-17|x| (8 significant bits per
    plane, rounding to nearest: |x - h0| < 2^-8 |x| puts x - h0 at least one
    binade lower), each reached within a factor 2 on random data; three planes
    hold a normal float32 exactly (the last residual has at most 8 significant
    bits); every plane holds bf16 values."""
    g = orc.rng(96)
    X = orc.wide_range(g, (64, 256), emin=-30, emax=30)
    x = X.astype(np.float64)
    h = orc.bf16_split(X, 3)
    for p in h:
        assert np.all((p.astype(np.float32).view(np.uint32) & 0xFFFF) == 0)
    for n, lim in ((1, 2.0 ** -8), (2, 2.0 ** -17)):
        res = np.abs(x - sum(h[:n]))
        assert np.all(res <= lim * np.abs(x))
        assert np.max(res / np.abs(x)) > lim / 2
    assert np.array_equal(h[0] + h[1] + h[2], x)
    assert [len(orc.bf16_split(X, n)) for n in (1, 2, 3)] == [1, 2, 3]
    assert orc.bf16_products(1) == [(0, 0)]
    assert sorted(orc.bf16_products(2)) == [(0, 0), (0, 1), (1, 0)]
    assert sorted(orc.bf16_products(3)) == [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (2, 0)]


def test_bf16_product_exact_on_small_integers_and_catches_dropped_terms():
    g = orc.rng(97)
    A = g.integers(-128, 129, size=(9, 40)).astype(np.float32)     # 8 significant bits: h0 = x
    B = g.integers(-128, 129, size=(40, 11)).astype(np.float32)
    for NP in (1, 2, 3):
        assert np.array_equal(orc.bf16_product(A, B, NP), A.astype(np.float64) @ B.astype(np.float64))
    A = orc.uniform(g, (32, 512)); B = orc.uniform(g, (512, 24))
    exact = A.astype(np.float64) @ B.astype(np.float64)
    scale = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    # Per-product bounds of the dropped terms: BF16X3 3*2^-16, BF16X6 8*2^-24;
    # the next smaller product set misses each.
    assert np.all(np.abs(orc.bf16_product(A, B, 2) - exact) <= 3 * 2.0 ** -16 * scale)
    assert not np.all(np.abs(orc.bf16_product(A, B, 1) - exact) <= 3 * 2.0 ** -16 * scale)
    assert np.all(np.abs(orc.bf16_product(A, B, 3) - exact) <= 8 * 2.0 ** -24 * scale)
    assert not np.all(np.abs(orc.bf16_product(A, B, 2) - exact) <= 8 * 2.0 ** -24 * scale)


@pytest.mark.parametrize("NP,K,M", [(1, 5, 64), (1, 64, 64), (2, 5, 64), (2, 24, 64), (3, 1, 256)])
def test_bf16_emulation_check_rejects_a_truncating_split(NP, K, M):
    """The GPU comparison (tests/test_gpu_bf16_emulation.py) holds a result to
    the emulation within the FP32 accumulation bound at these K: a float32
    rounding of the emulation passes it, a split that truncates fails it."""
    g = orc.rng(98 + NP * 100 + K)
    A, B = orc.uniform(g, (M, K)), orc.uniform(g, (K, M))
    emu = orc.bf16_product(A, B, NP)
    scale = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    assert orc.check_bound(emu.astype(np.float32), emu, scale, orc.tolerance(K))[0]
    planted = _trunc_split_product(A, B, NP).astype(np.float32)
    assert not orc.check_bound(planted, emu, scale, orc.tolerance(K))[0]


# ---------------------------------------------------------------------------
# Convolution oracle (im2col + GEMM + bias + LeakyReLU, sgemm_gl1.h:166-218).

def _direct_conv(x, wts, k, pad, stride, bias, act, slope):
    """Direct convolution from the definition (no im2col): an independent
    formula for the same result."""
    ich, h, w = x.shape
    ch = wts.shape[0]
    hcol, wcol = (h + 2 * pad - k) // stride + 1, (w + 2 * pad - k) // stride + 1
    xp = np.zeros((ich, h + 2 * pad, w + 2 * pad))
    xp[:, pad:pad + h, pad:pad + w] = x
    wk = wts.reshape(ch, ich, k, k).astype(np.float64)
    out = np.zeros((ch, hcol, wcol))
    for i in range(k):
        for j in range(k):
            win = xp[:, i:i + stride * hcol:stride, j:j + stride * wcol:stride]     # ich x hcol x wcol
            out += np.einsum("oc,cyx->oyx", wk[:, :, i, j], win)
    out = out.reshape(ch, -1)
    if bias is not None:
        out += bias.astype(np.float64)[:, None]
    if act:
        out = np.where(out > 0, out, slope * out)
    return out


@pytest.mark.parametrize("ich,h,w,k,pad,stride,ch", [(1, 5, 5, 3, 1, 1, 1), (3, 7, 6, 3, 1, 2, 4),
                                                      (2, 9, 9, 5, 2, 1, 3), (4, 8, 5, 1, 0, 1, 2),
                                                      (2, 6, 7, 4, 2, 2, 5)])
def test_conv_oracle_matches_direct_convolution(ich, h, w, k, pad, stride, ch):
    g = orc.rng(h * 100 + w * 10 + k)
    x = orc.uniform(g, (ich, h, w))
    wts = orc.uniform(g, (ch, ich * k * k))
    bias = orc.uniform(g, (ch,))
    out, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias, act=1, slope=0.1)
    ref = _direct_conv(x.astype(np.float64), wts, k, pad, stride, bias, 1, 0.1)
    assert np.allclose(out, ref, rtol=0, atol=1e-12)
    assert np.all(scale >= np.abs(out) - 1e-12)
    # planted mistakes: a wrong pad and a wrong activation slope are caught
    bad = _direct_conv(x.astype(np.float64), wts, k, pad + 1, stride, bias, 1, 0.1) if k > 1 else None
    if bad is not None and bad.shape == ref.shape:
        assert not np.allclose(out, bad, atol=1e-6)
    assert not np.allclose(out, _direct_conv(x.astype(np.float64), wts, k, pad, stride, bias, 1, 0.2), atol=1e-6)


def test_im2col_known_answer():
    # 1 channel 3x3 image, 2x2 kernel, no pad, stride 1: 4 rows (kernel
    # offsets) x 4 columns (output pixels).
    x = np.arange(1, 10, dtype=np.float32).reshape(1, 3, 3)
    col = orc.im2col(x, 1, 3, 3, 2, 0, 1)
    assert col.tolist() == [[1, 2, 4, 5], [2, 3, 5, 6], [4, 5, 7, 8], [5, 6, 8, 9]]
    # pad 1, stride 2, 3x3 kernel: the corner taps of the first pixel are outside the image
    col = orc.im2col(x, 1, 3, 3, 3, 1, 2)
    assert col.shape == (9, 4) and col[0, 0] == 0 and col[4, 0] == 1 and col[8, 3] == 0 and col[4, 3] == 9


# ---------------------------------------------------------------------------
# BLAS-1/2 oracle vs the compiled reference's saxpy/daxpy/sgemv/dgemv.

def _blas_golden():
    import ast
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_blas12.npz"))
    return d, [ast.literal_eval(s) for s in d["index"]]


def test_blas12_oracle_matches_reference_golden():
    d, index = _blas_golden()
    assert len(index) == 224
    diverging = 0
    for (key, name, N, incx, incy, alpha, beta, trans, M, lda, xbug) in index:
        dt = np.float64 if name[0] == "d" else np.float32
        eps = orc.EPS64 if dt == np.float64 else orc.EPS32
        x, y0, yref = d[key + "_x"], d[key + "_y0"], d[key + "_y"]
        y = y0.copy()
        if "axpy" in name:
            orc.axpy(N, alpha, x, incx, y, incy)
            # y + alpha*x: one or two roundings (the reference may contract to an FMA)
            sc = np.abs(y0.astype(np.float64)).copy()
            sc[::incy][:N] += abs(alpha) * np.abs(x[::incx][:N].astype(np.float64))
            assert np.all(np.abs(y.astype(np.float64) - yref) <= 2 * eps * sc), key
            continue
        A = d[key + "_A"]
        orc.gemv(trans, M, N, alpha, A, lda, x, incx, beta, y, incy, x_by_incy=True)
        a, xv = orc.gemv_dense(trans, M, N, A, lda, x, incy)        # the reference's reading of x
        scale = abs(alpha) * (np.abs(a) @ np.abs(xv)) + abs(beta) * np.abs(y0[::incy][:M].astype(np.float64))
        tol = 2 * (N + 2) * eps * scale
        assert np.all(np.abs(y[::incy][:M] - yref[::incy][:M]) <= tol), key
        if xbug and N > 1 and alpha != 0:
            y2 = y0.copy()
            orc.gemv(trans, M, N, alpha, A, lda, x, incx, beta, y2, incy, x_by_incy=False)
            diverging += not np.allclose(y2, yref, rtol=1e-3)
    assert diverging > 0   # the reference's x-indexing (incy) really differs from incx


def _axpy3_golden():
    d = np.load(os.path.join(GOLDEN, "ref_axpy3.npz"))
    return d, [ast.literal_eval(t) for t in d["index"]]


def test_axpy3_oracle_matches_reference_golden_bitwise():
    """saxpy_avx / daxpy_avx (ugemm.h:41-74) as the reference's FMA build
    computes them: the oracle's one-rounding z = a*x + y equals every golden
    output bit for bit on the first N elements; the reference's whole-vector
    loop also wrote past N (the divergence the GPU entry avoids)."""
    d, index = _axpy3_golden()
    assert len(index) == 20
    for key, name, N, a in index:
        x, y, zfull = d[key + "_x"], d[key + "_y"], d[key + "_zfull"]
        z = orc.axpy3(N, a, x[:N], y[:N])
        assert np.array_equal(z.view(np.uint8), zfull[:N].view(np.uint8)), key
        if N == 13:
            assert np.all(zfull[13:] != 0), key     # written beyond N by the reference


def test_axpy3_golden_pins_the_fma_branch():
    """A planted mistake: the non-FMA evaluation (a*x rounded, then + y) differs
    from the golden outputs on some elements, so the fixture pins one rounding."""
    d, index = _axpy3_golden()
    for key, name, N, a in index:
        if N < 1000 or a != -0.7:     # (1.5 * a 24-bit input is exact in float64)
            continue
        x, y, zfull = d[key + "_x"], d[key + "_y"], d[key + "_zfull"]
        dt = x.dtype.type
        two_roundings = (dt(a) * x[:N]).astype(x.dtype) + y[:N]
        assert not np.array_equal(two_roundings, zfull[:N]), key


def test_blas12_oracle_matches_numpy():
    g = orc.rng(1212)
    for trans in "NT":
        M, N, pad, incx, incy = 37, 53, 2, 3, 2
        rows, cols = (M, N) if trans == "N" else (N, M)
        lda = rows + pad
        A = orc.uniform(g, (cols * lda,), dtype=np.float64)
        x = orc.uniform(g, ((N - 1) * incx + 1,), dtype=np.float64)
        y0 = orc.uniform(g, ((M - 1) * incy + 1,), dtype=np.float64)
        y = y0.copy()
        orc.gemv(trans, M, N, 1.5, A, lda, x, incx, 0.5, y, incy)
        a = A.reshape(cols, lda)[:, :rows]          # column-major rows x cols, stored by columns
        a = a.T if trans == "N" else a              # 'N': a(m, n) = A[m + n*lda]; 'T': A[n + m*lda]
        want = 1.5 * a @ x[::incx][:N] + 0.5 * y0[::incy][:M]
        assert np.allclose(y[::incy][:M], want, rtol=1e-13, atol=1e-13)
        assert np.array_equal(y[1::incy], y0[1::incy])          # gaps untouched


# ---------------------------------------------------------------------------
# Config 1: the reference's CPU paths at 1024^3 (row-major NN, alpha 1,
# beta 0, U[-1,1) from seed 0x55474D4D01, C zero) against FP64.

def test_config1_cpu_paths_at_1024():
    """sgemm_avx (sgemm_avx256.h:392-455) and sgemm_c (gemm_cpu.h:284-363),
    compiled from the reference sources, and the oracle's restatement of the
    gemm_cpu blocking, all within the plain FP32 bound of numpy's FP64
    product; a planted wrong beta is caught by the same check."""
    n = 1024
    g = orc.rng(0x55474D4D01)
    A, B = orc.uniform(g, (n, n)), orc.uniform(g, (n, n))
    exact = A.astype(np.float64) @ B.astype(np.float64)
    scale = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    tol = orc.tolerance(n)
    runs = {"oracle_sgemm_packed": lambda C: orc.sgemm_packed("R", "N", "N", n, n, n, 1.0, A, n, B, n, 0.0, C, n)}
    L = orc.ref_lib()
    if L is not None:
        for name in ("sgemm_avx", "sgemm_c"):
            fn = getattr(L, name)
            runs[name] = (lambda f: lambda C: f(b"R", b"N", b"N", n, n, n, 1.0, A.ctypes.data, n, B.ctypes.data, n,
                                                0.0, C.ctypes.data, n))(fn)
    for name, run in runs.items():
        C = np.zeros((n, n), np.float32)
        run(C)
        ok, worst = orc.check_bound(C, exact, scale, tol)
        assert ok, f"{name}: worst/bound {worst:.3g}"
    C = np.ones((n, n), np.float32)          # beta = 0 must not keep C: a stale C is far outside the bound
    ok, _ = orc.check_bound(C + exact.astype(np.float32), exact, scale, tol)
    assert not ok
