"""Grouped SGEMM (one persistent launch over many uneven problems, optionally
replayed as a hipGraph): every problem's C must meet the FP32 bound against
the oracle, replays must be bitwise identical, tuning must keep results."""
import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case

pytestmark = pytest.mark.gpu


class DevCase:
    def __init__(self, case):
        self.case = case
        self.dA = hip.DeviceArray.from_host(case.A)
        self.dB = hip.DeviceArray.from_host(case.B)
        self.dC = hip.DeviceArray.from_host(case.C0)

    def problem(self):
        c = self.case
        return ug.Problem(c.ta.encode(), c.tb.encode(), c.M, c.N, c.K, c.alpha, c.beta, self.dA.ptr, c.lda,
                          self.dB.ptr, c.ldb, self.dC.ptr, c.ldc)

    def reset(self):
        self.dC.upload(self.case.C0)

    def check(self):
        C = self.dC.download()
        return self.case.judge(C)


def uneven_cases(n, seed, max_cin=64, sizes=(7, 14, 28)):
    # im2col-shaped problems (the conv -> GEMM mapping of sgemm_ocl1.h:271-300):
    # M = Cout, N = H*W, K = Cin*3*3, mixed layouts, alpha/beta and ld padding.
    g = orc.rng(seed)
    out = []
    for i in range(n):
        M = 32 * int(g.integers(1, 9))
        hw = int(g.choice(sizes))
        N = hw * hw
        K = 9 * 16 * int(g.integers(1, max_cin // 16 + 1))
        ta, tb = ("N", "T")[i % 2], ("N", "T")[(i // 2) % 2]
        alpha, beta = ((1.0, 0.0), (1.5, 0.5), (-1.0, 1.0))[i % 3]
        out.append(Case("R", ta, tb, M, N, K, alpha, beta, seed=seed * 1000 + i, pad=(i % 3)))
    # edge problems: tiny, K = 0, alpha = 0
    out.append(Case("R", "N", "N", 1, 1, 1, 1.0, 0.0, seed=seed + 7))
    out.append(Case("R", "N", "N", 33, 17, 0, 1.0, 0.5, seed=seed + 8))
    out.append(Case("R", "T", "N", 40, 50, 60, 0.0, 0.25, seed=seed + 9))
    return out


@pytest.mark.parametrize("capture", [False, True])
def test_grouped_matches_oracle(capture):
    cases = [DevCase(c) for c in uneven_cases(40, 5)]
    plan = ug.GroupedPlan([d.problem() for d in cases], capture=capture)
    s = hip.Stream()
    plan.run(s)
    s.synchronize()
    for d in cases:
        ok, worst = d.check()
        assert ok, (d.case.M, d.case.N, d.case.K, d.case.ta, d.case.tb, worst)
    plan.destroy()
    s.destroy()


def test_graph_replays_are_identical():
    cases = [DevCase(c) for c in uneven_cases(24, 9)]
    for d in cases:
        d.case.beta = 0.0
    plan = ug.GroupedPlan([d.problem() for d in cases], capture=True)
    s = hip.Stream()
    first = None
    for r in range(3):
        plan.run(s)
        s.synchronize()
        now = [d.dC.download() for d in cases]
        if first is None:
            first = now
        else:
            for a, b in zip(first, now):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        # eager work between replays
        ug.sgemm_dev("R", "N", "N", 64, 64, 64, 1.0, cases[0].dA, cases[0].case.lda, cases[0].dB,
                     cases[0].case.ldb, 0.0, hip.DeviceArray((64, 64), np.float32), 64, stream=s) \
            if cases[0].case.ta == "N" and cases[0].case.lda >= 64 and cases[0].case.M >= 64 else None
        s.synchronize()
    for d in cases:
        ok, worst = d.check()
        assert ok, worst
    plan.destroy()
    s.destroy()


def test_grouped_tuning_keeps_results():
    cases = [DevCase(c) for c in uneven_cases(16, 11)]
    plan = ug.GroupedPlan([d.problem() for d in cases], capture=True)
    v = plan.tune()
    assert 0 <= v < len(ug.grouped_variants())
    for d in cases:
        d.reset()
    s = hip.Stream()
    plan.run(s)
    s.synchronize()
    for d in cases:
        assert d.check()[0]
    plan.destroy()
    s.destroy()


@pytest.mark.parametrize("mode", [ug.MODE_FP32, ug.MODE_BF16X3])
def test_grouped_tuning_leaves_c_untouched(mode):
    """ugemm_hip_tune_grouped runs the plan many times; with beta != 0 that
    would leave C as neither the input nor one run's result, so the tuner
    restores every C: after tune() C still holds C0, and one run then meets
    the bound without re-uploading C.  C0 is uploaded on another stream
    without a wait: the tuner waits for the device first."""
    cases = [DevCase(c) for c in uneven_cases(12, 21)]
    s = hip.Stream()
    for d in cases:
        hip.check(hip.lib().hipMemcpyAsync(d.dC.ptr, d.case.C0.ctypes.data, d.dC.nbytes, hip.H2D, s.handle),
                  "async upload")
    plan = ug.GroupedPlan([d.problem() for d in cases], mode=mode, capture=True)
    plan.tune()
    for d in cases:
        assert np.array_equal(d.dC.download(), d.case.C0)
    plan.run(s)
    s.synchronize()
    for d in cases:
        ok, worst = d.case.judge(d.dC.download(), mode=mode)
        assert ok, worst
    plan.destroy()
    s.destroy()


def test_grouped_mixed_tile_shapes():
    """Variant 4 gives each problem its own tile shape (128 x 128 or 64 x 128
    by the padding of its rows): one plan mixing both, every layout, checked
    per problem, replayed from a graph twice (bitwise equal)."""
    assert ug.grouped_variants()[4].startswith("fp32_mixed")
    assert len(ug.grouped_variants()) == 5
    cases = []
    for i, M in enumerate((32, 64, 96, 128, 160, 192, 224, 256, 300, 352, 512, 33)):
        ta, tb = ("N", "T")[i % 2], ("N", "T")[(i // 2) % 2]
        cases.append(DevCase(Case("R", ta, tb, M, 49 + 37 * i, 144 * (1 + i % 4), 1.5, 0.5, seed=400 + i, pad=i % 3)))
    ug.force_variant(ug.MODE_FP32, 4)
    try:
        plan = ug.GroupedPlan([d.problem() for d in cases], capture=True)
    finally:
        ug.force_variant(ug.MODE_FP32, -1)
    outs = []
    for _ in range(2):
        for d in cases:
            d.reset()
        plan.run(None)
        hip.synchronize()
        outs.append([d.dC.download() for d in cases])
        for d in cases:
            ok, worst = d.check()
            assert ok, f"M={d.case.M}: {worst:.3g}"
    plan.destroy()
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


SPLIT_GROUPED = [ug.MODE_BF16X1, ug.MODE_BF16X3, ug.MODE_BF16X6, ug.MODE_FP16X3]


@pytest.mark.parametrize("mode", SPLIT_GROUPED)
@pytest.mark.parametrize("variant", [-1, 0, 1, 2, 3, 4, 5, 6])
def test_grouped_split_forms(mode, variant):
    """Grouped plans in the split forms: every problem's planes from one
    preparation launch (FP16X3: after one row-maximum launch), one split-core
    block per tile; uneven im2col-shaped problems of every layout (and the
    tiny / K = 0 / alpha = 0 edge problems) meet the form's bound
    (helpers.MODE_EXTRA); -1 = default tile."""
    cases = [DevCase(c) for c in uneven_cases(14, 50 + variant + 7 * mode)]
    ug.force_variant(mode, variant)
    try:
        plan = ug.GroupedPlan([d.problem() for d in cases], mode=mode, capture=False)
    finally:
        ug.force_variant(mode, -1)
    plan.run(None)
    hip.synchronize()
    for d in cases:
        ok, worst = d.case.judge(d.dC.download(), mode=mode)
        assert ok, f"mode {mode} variant {variant} {d.case.M}x{d.case.N}x{d.case.K}: {worst:.3g}"
    plan.destroy()


def uniform_b_cases(n, seed, tb, sizes=(7, 14, 28)):
    """im2col-shaped problems whose B all have layout tb (A mixed), with
    padded and odd leading dimensions, plus tiny, K = 0 and alpha = 0 ones."""
    g = orc.rng(seed)
    out = []
    for i in range(n):
        M = 32 * int(g.integers(1, 9))
        hw = int(g.choice(sizes))
        K = 9 * 16 * int(g.integers(1, 5))
        alpha, beta = ((1.0, 0.0), (1.5, 0.5), (-1.0, 1.0))[i % 3]
        out.append(Case("R", ("N", "T")[i % 2], tb, M, hw * hw, K, alpha, beta, seed=seed * 1000 + i, pad=(i % 3)))
    out.append(Case("R", "N", tb, 1, 1, 1, 1.0, 0.0, seed=seed + 7))
    out.append(Case("R", "T", tb, 33, 17, 0, 1.0, 0.5, seed=seed + 8))
    out.append(Case("R", "N", tb, 40, 50, 60, 0.0, 0.25, seed=seed + 9))
    out.append(Case("R", "T", tb, 129, 131, 77, 1.5, 0.5, seed=seed + 10, pad=1))
    return out


@pytest.mark.parametrize("mode", SPLIT_GROUPED)
@pytest.mark.parametrize("variant", [8, 9, 10])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_grouped_split_forms_split_b(mode, variant, tb):
    """Grouped-only variants that split B in the core from its FP32 source
    (no B planes): every B of the plan has one layout; the form's bound; a
    captured plan replays bitwise identically."""
    assert len(ug.grouped_variants(mode)) == 11 and ug.grouped_variants(mode)[variant].endswith("_split_b")
    cases = [DevCase(c) for c in uniform_b_cases(12, 70 + variant + 3 * mode + (tb == "T"), tb)]
    ug.force_variant(mode, variant)
    try:
        plan = ug.GroupedPlan([d.problem() for d in cases], mode=mode, capture=True)
    finally:
        ug.force_variant(mode, -1)
    outs = []
    for _ in range(2):
        for d in cases:
            d.reset()
        plan.run(None)
        hip.synchronize()
        outs.append([d.dC.download() for d in cases])
    for d, C in zip(cases, outs[0]):
        ok, worst = d.case.judge(C, mode=mode)
        assert ok, f"mode {mode} variant {variant} {d.case.ta}{tb} {d.case.M}x{d.case.N}x{d.case.K}: {worst:.3g}"
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    plan.destroy()


def test_grouped_split_b_needs_one_b_layout():
    """A split-B variant on a plan whose B layouts differ is refused
    (UNSUPPORTED); the tuner skips such variants."""
    cases = [DevCase(c) for c in uneven_cases(6, 88)]
    split_b = len(ug.variants(ug.MODE_BF16X3))   # the first split-B variant
    ug.force_variant(ug.MODE_BF16X3, split_b)
    try:
        with pytest.raises(ug.UgemmError) as e:
            ug.GroupedPlan([d.problem() for d in cases], mode=ug.MODE_BF16X3, capture=False)
        assert e.value.status == ug.ERR_UNSUPPORTED
    finally:
        ug.force_variant(ug.MODE_BF16X3, -1)
    plan = ug.GroupedPlan([d.problem() for d in cases], mode=ug.MODE_BF16X3, capture=False)
    assert plan.tune() < split_b
    plan.destroy()


def test_grouped_split_form_graph_replays_and_tuning(tmp_path):
    """A captured BF16x3 plan replays bitwise identically; tuning picks one
    of the split-core variants and leaves results within the bound."""
    mode = ug.MODE_BF16X3
    cases = [DevCase(c) for c in uneven_cases(20, 91, max_cin=128, sizes=(7, 14, 28, 56))]
    plan = ug.GroupedPlan([d.problem() for d in cases], mode=mode, capture=True)
    v = plan.tune()
    assert 0 <= v < len(ug.grouped_variants(mode))
    outs = []
    for _ in range(2):
        for d in cases:
            d.reset()
        plan.run(None)
        hip.synchronize()
        outs.append([d.dC.download() for d in cases])
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for d, C in zip(cases, outs[0]):
        ok, worst = d.case.judge(C, mode=mode)
        assert ok, worst
    plan.destroy()
    # The tuned split-core variant (up to 6) survives a save / load round trip.
    L = ug.load()
    path = str(tmp_path / "tune.txt").encode()
    assert L.ugemm_hip_tune_save(path) == ug.OK
    rows = [l.split() for l in open(path.decode()).read().split("\n")[1:] if l.strip()]
    assert L.ugemm_hip_tune_load(path) == len(rows)
    assert any(int(r[0]) == 16 and int(r[2]) == mode and int(r[6]) == v for r in rows)   # 16: grouped key


@pytest.mark.parametrize("variant", [-1, 0, 1, 2, 3, 4])
def test_grouped_fp32_mfma(variant):
    """FP32-MFMA plans: the grouped persistent kernel with FP32 matrix-core
    tiles (variant 4: 128x128 or 64x128 per problem); every problem of the
    uneven batch (all layouts, unaligned leading dimensions, the tiny / K = 0
    / alpha = 0 edge problems) meets the plain FP32 bound."""
    mode = ug.MODE_FP32_MFMA
    cases = [DevCase(c) for c in uneven_cases(16, 70 + variant)]
    ug.force_variant(mode, variant)
    try:
        plan = ug.GroupedPlan([d.problem() for d in cases], mode=mode, capture=False)
    finally:
        ug.force_variant(mode, -1)
    plan.run(None)
    hip.synchronize()
    for d in cases:
        ok, worst = d.case.judge(d.dC.download(), mode=mode)
        assert ok, f"variant {variant} {d.case.M}x{d.case.N}x{d.case.K} {d.case.ta}{d.case.tb}: {worst:.3g}"
    plan.destroy()


def test_grouped_fp32_mfma_graph_replays_tuning_and_cache(tmp_path):
    """A captured FP32-MFMA plan replays bitwise identically after tuning;
    the tuning cache keeps the grouped entries of every form on reload."""
    mode = ug.MODE_FP32_MFMA
    names = ug.grouped_variants(mode)
    assert len(names) == 5 and all(n.startswith("f32_mfma_") for n in names)
    cases = [DevCase(c) for c in uneven_cases(20, 93, max_cin=128, sizes=(7, 14, 28, 56))]
    plan = ug.GroupedPlan([d.problem() for d in cases], mode=mode, capture=True)
    v = plan.tune()
    assert 0 <= v < 5
    outs = []
    for _ in range(3):
        for d in cases:
            d.reset()
        plan.run(None)
        hip.synchronize()
        outs.append([d.dC.download() for d in cases])
    for run in outs[1:]:
        for a, b in zip(outs[0], run):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for d, C in zip(cases, outs[0]):
        ok, worst = d.case.judge(C, mode=mode)
        assert ok, worst
    plan.destroy()
    L = ug.load()
    path = str(tmp_path / "tune.txt").encode()
    assert L.ugemm_hip_tune_save(path) == ug.OK
    rows = [l for l in open(path.decode()).read().split("\n")[1:] if l.strip()]
    assert L.ugemm_hip_tune_load(path) == len(rows)


def test_grouped_mode_variant_tables():
    L = ug.load()
    assert ug.grouped_variants(ug.MODE_FP32) == ug.grouped_variants()
    for mode in (ug.MODE_BF16X1, ug.MODE_BF16X3, ug.MODE_BF16X6, ug.MODE_FP16X3):
        # the single-GEMM split-core table, then the grouped-only split-B variants
        g, v = ug.grouped_variants(mode), ug.variants(mode)
        assert g[:len(v)] == v and len(g) == len(v) + 3
        assert g[len(v)] == v[1] + "_split_b" and g[len(v) + 2] == v[7] + "_split_b"
    for mode in (ug.MODE_FP8_BLOCK, 42, -1):
        assert L.ugemm_hip_grouped_mode_variant_count(mode) == ug.ERR_ARG
    assert L.ugemm_hip_grouped_mode_variant_name(ug.MODE_FP32_MFMA, 5) is None


# A forced single-GEMM variant the grouped kernel lacks (5) is ignored.
@pytest.mark.parametrize("variant", range(6))
def test_grouped_every_variant(variant):
    ug.force_variant(ug.MODE_FP32, variant)
    try:
        cases = [DevCase(c) for c in uneven_cases(12, 13 + variant)]
        ug.load().ugemm_hip_sgemm_grouped(len(cases), (ug.Problem * len(cases))(*[d.problem() for d in cases]),
                                          ug.MODE_FP32, None)
        hip.synchronize()
        for d in cases:
            assert d.check()[0]
    finally:
        ug.force_variant(ug.MODE_FP32, -1)


def test_grouped_argument_errors():
    L = ug.load()
    bad = ug.Problem(b"X", b"N", 4, 4, 4, 1.0, 0.0, 1, 4, 1, 4, 1, 4)
    arr = (ug.Problem * 1)(bad)
    assert L.ugemm_hip_sgemm_grouped(1, arr, ug.MODE_FP32, None) == ug.ERR_ARG
    bad = ug.Problem(b"N", b"N", 4, 4, 4, 1.0, 0.0, None, 4, None, 4, None, 4)
    arr = (ug.Problem * 1)(bad)
    assert L.ugemm_hip_sgemm_grouped(1, arr, ug.MODE_FP32, None) == ug.ERR_ARG
    ok = ug.Problem(b"N", b"N", 4, 4, 4, 1.0, 0.0, 16, 4, 16, 4, 16, 4)
    arr = (ug.Problem * 1)(ok)
    # The form without a grouped path is refused before any work (the
    # pointers here are not device memory).
    assert L.ugemm_hip_sgemm_grouped(1, arr, ug.MODE_FP8_BLOCK, None) == ug.ERR_UNSUPPORTED
    assert L.ugemm_hip_sgemm_grouped(1, arr, 42, None) == ug.ERR_UNSUPPORTED
    assert L.ugemm_hip_sgemm_grouped(0, None, ug.MODE_FP32, None) == ug.OK


@pytest.mark.parametrize("capture", [False, True])
def test_grouped_fewer_tiles_than_queues(capture):
    """Three one-tile problems: a grid of 3 workgroups over 8 queues, five of
    them empty; every workgroup must still find its tile and exit."""
    cases = [DevCase(Case("R", "N", "N", 5 + i, 7 + 2 * i, 9 + i, 1.5, 0.5, seed=300 + i)) for i in range(3)]
    plan = ug.GroupedPlan([d.problem() for d in cases], capture=capture)
    s = hip.Stream()
    for _ in range(2):
        for d in cases:
            d.reset()
        plan.run(s)
        s.synchronize()
        for d in cases:
            ok, worst = d.check()
            assert ok, worst
    plan.destroy()
    s.destroy()


def test_grouped_more_tiles_than_workgroups_two_plans():
    """More tiles than resident workgroups (each takes several, and queues are
    drained by neighbours), with two plans alternating on one stream: each
    plan's queue heads are cleared before each of its launches."""
    ug.force_variant(ug.MODE_FP32, 1)     # 64 x 64 tiles: 16 per problem
    try:
        a = [DevCase(Case("R", "N", "N", 256, 256, 144, 1.0, 0.0, seed=400 + i)) for i in range(120)]
        b = [DevCase(Case("R", "T", "N", 96, 200, 70, -1.0, 1.0, seed=600 + i, pad=1)) for i in range(20)]
        pa = ug.GroupedPlan([d.problem() for d in a], capture=True)
        pb = ug.GroupedPlan([d.problem() for d in b], capture=False)
    finally:
        ug.force_variant(ug.MODE_FP32, -1)
    s = hip.Stream()
    for r in range(2):
        for d in a + b:
            d.reset()
        pa.run(s)
        pb.run(s)
        pa.run(s)     # beta == 0 for plan a: a second run gives the same C
        s.synchronize()
        for d in a + b:
            ok, worst = d.check()
            assert ok, (r, d.case.M, d.case.N, d.case.K, worst)
    pa.destroy()
    pb.destroy()
    s.destroy()


@pytest.mark.parametrize("variant", [1, 8, 10])
def test_grouped_fp16x3_row_scaled_preparation_at_k_limit(variant):
    """FP16X3 plans whose k-contiguous A (K padded to at most 8192) gets its
    row maxima in its own preparation (two rows per block), beside problems
    that take the row-maximum pass (K = 8200, and A transposed), against the
    exact emulation of each product; plane-reading and split-B variants."""
    mode = ug.MODE_FP16X3
    specs = [("N", 66, 40, 8192), ("N", 130, 50, 8200), ("T", 70, 33, 300), ("N", 3, 7, 8190)]
    cases = [DevCase(Case("R", ta, "N", M, N, K, 1.0, 0.0, seed=1000 + i, dist="wide"))
             for i, (ta, M, N, K) in enumerate(specs)]
    ug.force_variant(mode, variant)
    try:
        plan = ug.GroupedPlan([d.problem() for d in cases], mode=mode, capture=True)
    finally:
        ug.force_variant(mode, -1)
    plan.run(None)
    hip.synchronize()
    for d in cases:
        c = d.case
        C = d.dC.download()
        Aop = orc.op_matrix("R", c.ta, c.A, c.M, c.K)
        Bop = orc.op_matrix("R", c.tb, c.B, c.K, c.N)
        emu = orc.fp16x3_product(Aop, Bop)
        scale = np.abs(Aop.astype(np.float64)) @ np.abs(Bop.astype(np.float64))
        ok, worst = orc.check_bound(C[:, :c.N], emu, scale, orc.tolerance(c.K))
        assert ok, f"variant {variant} {c.ta}N {c.M}x{c.N}x{c.K}: worst/bound {worst:.3g}"
    plan.destroy()
