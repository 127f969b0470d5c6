"""Fused bias + LeakyReLU epilogue and implicit-GEMM convolution vs the oracle.

The convolution oracle restates gl_convolution_LReLU (sgemm_gl1.h:191-218):
im2col (sgemm_gl1.h:166-190), the weights x col product, out += bias[c],
out = out > 0 ? out : 0.1 * out.  Acceptance: the FP32 elementwise bound
2(K+1) eps32 * scale with scale = sum |w||col| + |bias| (the activation only
shrinks errors).  The GEMM epilogue is checked against the oracle's exact
product followed by the same bias/activation."""
import ctypes
import zlib

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc
from tests.helpers import Case, MODE_EXTRA

pytestmark = pytest.mark.gpu

CONV_SHAPES = [
    # ich, h, w, k, pad, stride, ch
    (1, 5, 5, 3, 1, 1, 1),
    (3, 17, 13, 3, 1, 1, 16),
    (16, 14, 14, 3, 1, 1, 32),
    (8, 28, 28, 3, 1, 2, 24),
    (4, 9, 11, 5, 2, 1, 7),
    (32, 7, 7, 1, 0, 1, 64),
    (3, 33, 31, 4, 2, 2, 5),
    (64, 14, 14, 3, 1, 1, 130),
]


def _conv_dev(x, wts, bias, ich, h, w, k, pad, stride, ch, act, slope, variant=None, mode=ug.MODE_FP32):
    hcol, wcol = orc.conv_out_shape(h, w, k, pad, stride)
    d_x, d_w = hip.DeviceArray.from_host(x), hip.DeviceArray.from_host(wts)
    d_b = hip.DeviceArray.from_host(bias) if bias is not None else None
    d_o = hip.DeviceArray.from_host(np.full((ch, hcol * wcol), np.nan, np.float32))
    if variant is not None:
        ug.force_variant(mode, variant)
    try:
        ug.conv2d_dev(d_x, ich, w, h, d_w, k, pad, stride, d_o, ch, bias=d_b, act=act, slope=slope, mode=mode)
    finally:
        if variant is not None:
            ug.force_variant(mode, -1)
    hip.synchronize()
    return d_o.download()


@pytest.mark.parametrize("shape", CONV_SHAPES)
@pytest.mark.parametrize("act,with_bias", [(ug.ACT_LEAKY_RELU, True), (ug.ACT_NONE, False)])
def test_conv2d_matches_oracle(shape, act, with_bias):
    ich, h, w, k, pad, stride, ch = shape
    g = orc.rng(hash(shape) % (1 << 31))
    x = orc.uniform(g, (ich, h, w))
    wts = orc.uniform(g, (ch, ich * k * k))
    bias = orc.uniform(g, (ch,)) if with_bias else None
    print("conv", shape, act, with_bias, flush=True)
    got = _conv_dev(x, wts, bias, ich, h, w, k, pad, stride, ch, act, 0.1)
    exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias, act=act, slope=0.1)
    ok, worst = orc.check_bound(got, exact, scale, orc.tolerance(ich * k * k))
    assert ok, f"{shape}: worst/bound {worst:.3g}"


# Variants 4 and 5 (the larger single-GEMM tiles) fall back to variant 0.
@pytest.mark.parametrize("variant", range(6))
def test_conv2d_every_tile_variant(variant):
    ich, h, w, k, pad, stride, ch = 6, 15, 12, 3, 1, 1, 70
    g = orc.rng(700 + variant)
    x = orc.uniform(g, (ich, h, w))
    wts = orc.uniform(g, (ch, ich * k * k))
    bias = orc.uniform(g, (ch,))
    got = _conv_dev(x, wts, bias, ich, h, w, k, pad, stride, ch, ug.ACT_LEAKY_RELU, 0.1, variant=variant)
    exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias)
    assert orc.check_bound(got, exact, scale, orc.tolerance(ich * k * k))[0]


def test_convolution_LReLU_host_entry():
    """The gl_convolution_LReLU argument list, host arrays, slope 0.1."""
    ich, h, w, k, pad, stride, ch = 3, 20, 18, 3, 1, 1, 12
    g = orc.rng(801)
    x = orc.uniform(g, (ich, h, w))
    wts = orc.uniform(g, (ch, ich * k * k))
    bias = orc.uniform(g, (ch,))
    hcol, wcol = orc.conv_out_shape(h, w, k, pad, stride)
    out = np.zeros((ch, hcol * wcol), np.float32)
    ug.convolution_LReLU(x, ich, w, h, wts, k, pad, stride, out, ch, bias)
    exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias)
    assert orc.check_bound(out, exact, scale, orc.tolerance(ich * k * k))[0]
    assert np.any(out < 0) and np.any(out > 0)      # both branches of the activation


def test_conv2d_argument_errors():
    L = ug.load()
    d = hip.DeviceArray(64, np.float32)
    assert L.ugemm_hip_conv2d_dev(d.ptr, 1, 4, 4, d.ptr, 5, 0, 1, d.ptr, 1, None, 1, 0.1, 0, None) == ug.ERR_ARG
    assert L.ugemm_hip_conv2d_dev(d.ptr, 1, 4, 4, d.ptr, 3, 0, 0, d.ptr, 1, None, 1, 0.1, 0, None) == ug.ERR_ARG
    assert L.ugemm_hip_conv2d_dev(d.ptr, 1, 4, 4, d.ptr, 3, 0, 1, d.ptr, 1, None, 7, 0.1, 0, None) == ug.ERR_ARG
    assert L.ugemm_hip_conv2d_dev(d.ptr, 1, 4, 4, d.ptr, 3, 0, 1, d.ptr, 1, None, 1, 0.1,
                                  ug.MODE_FP8_BLOCK, None) == ug.ERR_UNSUPPORTED
    assert L.ugemm_hip_conv2d_dev(d.ptr, 1, 4, 4, d.ptr, 3, 0, 1, d.ptr, 1, None, 1, 0.1, 42, None) == ug.ERR_ARG


@pytest.mark.parametrize("shape", CONV_SHAPES)
@pytest.mark.parametrize("variant", [None, 0, 1, 3])
def test_conv2d_fp32_mfma(shape, variant):
    """Convolutions on the FP32 matrix cores: the im2col operand gathered in
    the FP32-MFMA tile loader, bias and LeakyReLU in its store; plain FP32
    bound.  Weights with K = ich*k*k not a multiple of 4 take scalar loads."""
    ich, h, w, k, pad, stride, ch = shape
    g = orc.rng(zlib.crc32(f"m32conv{shape}".encode()))
    x = orc.uniform(g, (ich, h, w))
    wts = orc.uniform(g, (ch, ich * k * k))
    bias = orc.uniform(g, (ch,))
    got = _conv_dev(x, wts, bias, ich, h, w, k, pad, stride, ch, ug.ACT_LEAKY_RELU, 0.1, variant=variant,
                    mode=ug.MODE_FP32_MFMA)
    exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias, act=1, slope=0.1)
    ok, worst = orc.check_bound(got, exact, scale, orc.tolerance(ich * k * k))
    assert ok, f"{shape} variant {variant}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("variant", range(6))
def test_conv2d_fp32_mfma_every_variant(variant):
    ich, h, w, k, pad, stride, ch = 24, 20, 21, 3, 1, 1, 70
    g = orc.rng(77 + variant)
    x, wts = orc.uniform(g, (ich, h, w)), orc.uniform(g, (ch, ich * k * k))
    got = _conv_dev(x, wts, None, ich, h, w, k, pad, stride, ch, ug.ACT_NONE, 0.1, variant=variant,
                    mode=ug.MODE_FP32_MFMA)
    exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, None, act=0, slope=0.1)
    ok, worst = orc.check_bound(got, exact, scale, orc.tolerance(ich * k * k))
    assert ok, f"variant {variant}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("mode", [ug.MODE_FP32, ug.MODE_BF16X3, ug.MODE_FP32_MFMA])
@pytest.mark.parametrize("splits", [1, 3, 0])
def test_conv2d_split_k(mode, splits):
    """A convolution with few output tiles and a long K (ich * k * k = 4608)
    split along K (forced 3, or automatic) matches the unsplit bound."""
    ich, h, w, k, pad, stride, ch = 512, 12, 12, 3, 1, 1, 40
    g = orc.rng(900 + splits + mode)
    x, wts, bias = orc.uniform(g, (ich, h, w)), orc.uniform(g, (ch, ich * k * k)), orc.uniform(g, (ch,))
    hcol, wcol = orc.conv_out_shape(h, w, k, pad, stride)
    dx, dw, db = hip.DeviceArray.from_host(x), hip.DeviceArray.from_host(wts), hip.DeviceArray.from_host(bias)
    dout = hip.DeviceArray((ch, hcol * wcol), np.float32)
    ug.force_splits(mode, splits)
    try:
        ug.conv2d_dev(dx, ich, w, h, dw, k, pad, stride, dout, ch, db, act=ug.ACT_LEAKY_RELU, slope=0.1, mode=mode)
        hip.synchronize()
    finally:
        ug.force_splits(mode, 0)
    exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias, act=1, slope=0.1)
    ok, worst = orc.check_bound(dout.download(), exact, scale,
                                orc.tolerance(ich * k * k, extra=MODE_EXTRA.get(mode, 0.0)))
    assert ok, f"mode {mode} splits {splits}: worst/bound {worst:.3g}"


@pytest.mark.parametrize("mode", [ug.MODE_BF16X3, ug.MODE_FP16X3, ug.MODE_BF16X1, ug.MODE_BF16X6])
@pytest.mark.parametrize("shape", [(3, 17, 13, 3, 1, 1, 40), (16, 28, 28, 3, 1, 1, 64), (5, 20, 31, 5, 2, 2, 9),
                                   (1, 9, 9, 1, 0, 1, 300), (32, 14, 14, 3, 1, 1, 300)])
def test_conv2d_split_forms(mode, shape):
    """Convolutions in the split forms: the im2col operand gathered into
    its planes, then the split core with bias and LeakyReLU; bound of the
    form (helpers.MODE_EXTRA) around the FP64 convolution."""
    ich, h, w, k, pad, stride, ch = shape
    g = orc.rng(zlib.crc32(f"{mode}{shape}".encode()))
    x = orc.uniform(g, (ich, h, w))
    wts = orc.uniform(g, (ch, ich * k * k))
    bias = orc.uniform(g, (ch,))
    hcol, wcol = orc.conv_out_shape(h, w, k, pad, stride)
    dx, dw, db = hip.DeviceArray.from_host(x), hip.DeviceArray.from_host(wts), hip.DeviceArray.from_host(bias)
    dout = hip.DeviceArray.from_host(np.full((ch, hcol * wcol), np.nan, np.float32))
    ug.conv2d_dev(dx, ich, w, h, dw, k, pad, stride, dout, ch, db, act=ug.ACT_LEAKY_RELU, slope=0.1, mode=mode)
    hip.synchronize()
    exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias, act=1, slope=0.1)
    ok, worst = orc.check_bound(dout.download(), exact, scale, orc.tolerance(ich * k * k, extra=MODE_EXTRA[mode]))
    assert ok, f"mode {mode} {shape}: worst/bound {worst:.3g}"


# ---------------------------------------------------------------------------
# GEMM with the fused epilogue.

def _epilogue(v, bias, bias_dim, act, slope, major):
    """The epilogue applied to a logical [m, n] result (either major)."""
    out = v.copy()
    if bias is not None:
        b = bias.astype(np.float64)
        out += b[:, None] if bias_dim == ug.BIAS_ROWS else b[None, :]
    if act == ug.ACT_LEAKY_RELU:
        out = np.where(out > 0, out, slope * out)
    return out


@pytest.mark.parametrize("mode", [ug.MODE_FP32, ug.MODE_BF16X3, ug.MODE_FP16X3, ug.MODE_BF16X6])
@pytest.mark.parametrize("major,ta,tb", [("R", "N", "N"), ("R", "T", "N"), ("C", "N", "T"), ("C", "T", "T")])
@pytest.mark.parametrize("bias_dim", [ug.BIAS_ROWS, ug.BIAS_COLS])
def test_sgemm_epilogue(mode, major, ta, tb, bias_dim):
    M, N, K = 150, 97, 133
    case = Case(major, ta, tb, M, N, K, 1.5, 0.5, seed=900 + mode * 10 + bias_dim, pad=1)
    g = orc.rng(950 + bias_dim)
    bias = orc.uniform(g, (M if bias_dim == ug.BIAS_ROWS else N,))
    dA, dB, dC = (hip.DeviceArray.from_host(case.A), hip.DeviceArray.from_host(case.B),
                  hip.DeviceArray.from_host(case.C0))
    d_b = hip.DeviceArray.from_host(bias)
    slope = 0.25
    ug.sgemm_ex_dev(major, ta, tb, M, N, K, case.alpha, dA, case.lda, dB, case.ldb, case.beta, dC, case.ldc,
                    mode=mode, bias=d_b, bias_dim=bias_dim, act=ug.ACT_LEAKY_RELU, slope=slope)
    hip.synchronize()
    got = case.logical(dC.download())
    exact, scale = case.exact()
    want = _epilogue(exact, bias, bias_dim, ug.ACT_LEAKY_RELU, slope, major)
    bscale = np.abs(bias.astype(np.float64))
    scale = scale + (bscale[:, None] if bias_dim == ug.BIAS_ROWS else bscale[None, :])
    ok, worst = orc.check_bound(got, want, scale, orc.tolerance(K, extra=MODE_EXTRA.get(mode, 0.0)))
    assert ok, f"mode {mode} {major}{ta}{tb} bias_dim {bias_dim}: {worst:.3g}"
    assert case.untouched_padding(dC.download())


def test_sgemm_epilogue_alpha_zero_and_activation_only():
    """alpha == 0 runs the scaling kernel, which applies the epilogue too;
    a NULL bias applies the activation alone."""
    M, N, K = 40, 30, 20
    case = Case("R", "N", "N", M, N, K, 0.0, 0.5, seed=990)
    bias = orc.uniform(orc.rng(991), (M,))
    dA, dB, dC = (hip.DeviceArray.from_host(case.A), hip.DeviceArray.from_host(case.B),
                  hip.DeviceArray.from_host(case.C0))
    ug.sgemm_ex_dev("R", "N", "N", M, N, K, 0.0, dA, K, dB, N, 0.5, dC, N, bias=hip.DeviceArray.from_host(bias),
                    act=ug.ACT_LEAKY_RELU, slope=0.1)
    hip.synchronize()
    want = _epilogue(0.5 * case.C0.astype(np.float64), bias, ug.BIAS_ROWS, ug.ACT_LEAKY_RELU, 0.1, "R")
    assert np.allclose(dC.download(), want, rtol=1e-6, atol=1e-7)
    dC.upload(case.C0)
    ug.sgemm_ex_dev("R", "N", "N", M, N, K, 1.0, dA, K, dB, N, 0.0, dC, N, act=ug.ACT_LEAKY_RELU, slope=0.1)
    hip.synchronize()
    exact = case.A.astype(np.float64) @ case.B.astype(np.float64)
    scale = np.abs(case.A.astype(np.float64)) @ np.abs(case.B.astype(np.float64))
    assert orc.check_bound(dC.download(), np.where(exact > 0, exact, 0.1 * exact), scale, orc.tolerance(K))[0]


def test_sgemm_epilogue_rejected_for_fp8():
    d = hip.DeviceArray(64, np.float32)
    L = ug.load()
    rc = L.ugemm_hip_sgemm_ex_dev(b"R", b"N", b"N", 4, 4, 4, 1.0, d.ptr, 4, d.ptr, 4, 0.0, d.ptr, 4,
                                  ug.MODE_FP8_BLOCK, d.ptr, 0, 1, 0.1, None)
    assert rc == ug.ERR_UNSUPPORTED
    rc = L.ugemm_hip_sgemm_ex_dev(b"R", b"N", b"N", 4, 4, 4, 1.0, d.ptr, 4, d.ptr, 4, 0.0, d.ptr, 4,
                                  ug.MODE_FP32, d.ptr, 5, 1, 0.1, None)
    assert rc == ug.ERR_ARG


# ---------------------------------------------------------------------------
# Grouped convolutions: one persistent launch over many independent layers.

def _conv_batch(n, seed):
    g = orc.rng(seed)
    out = []
    for i in range(n):
        ich = int(g.choice([1, 3, 8, 16, 24]))
        h = int(g.choice([5, 7, 13, 14, 20]))
        w = h + int(g.integers(0, 3))
        k = int(g.choice([1, 3, 3, 5]))
        pad = int(g.integers(0, k // 2 + 1))
        stride = int(g.choice([1, 1, 2]))
        ch = int(g.choice([4, 16, 33, 64]))
        x = orc.uniform(g, (ich, h, w))
        wts = orc.uniform(g, (ch, ich * k * k))
        bias = orc.uniform(g, (ch,)) if i % 3 else None
        act = ug.ACT_LEAKY_RELU if i % 2 else ug.ACT_NONE
        out.append((ich, h, w, k, pad, stride, ch, x, wts, bias, act))
    return out


def _conv_plan(layers, capture, mode=ug.MODE_FP32):
    keep, probs, outs = [], [], []
    for (ich, h, w, k, pad, stride, ch, x, wts, bias, act) in layers:
        hcol, wcol = orc.conv_out_shape(h, w, k, pad, stride)
        d_x, d_w = hip.DeviceArray.from_host(x), hip.DeviceArray.from_host(wts)
        d_b = hip.DeviceArray.from_host(bias) if bias is not None else None
        d_o = hip.DeviceArray.from_host(np.full((ch, hcol * wcol), np.nan, np.float32))
        keep += [d_x, d_w, d_b]
        outs.append(d_o)
        probs.append(ug.ConvProblem(d_x.ptr, ich, w, h, d_w.ptr, k, pad, stride, d_o.ptr, ch,
                                    d_b.ptr if d_b is not None else None, act, 0.1))
    return ug.GroupedPlan(probs, mode=mode, capture=capture), outs, keep


@pytest.mark.parametrize("capture", [False, True])
def test_grouped_convolutions_match_oracle(capture):
    layers = _conv_batch(24, 1234)
    plan, outs, keep = _conv_plan(layers, capture)
    s = hip.Stream()
    first = None
    for r in range(2):
        plan.run(s)
        s.synchronize()
        now = [o.download() for o in outs]
        if first is None:
            first = now
        else:
            assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(first, now))
    for (ich, h, w, k, pad, stride, ch, x, wts, bias, act), got in zip(layers, first):
        exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias, act=act, slope=0.1)
        ok, worst = orc.check_bound(got, exact, scale, orc.tolerance(ich * k * k))
        assert ok, ((ich, h, w, k, pad, stride, ch), worst)
    plan.destroy()
    s.destroy()


@pytest.mark.parametrize("variant", range(6))
def test_grouped_convolutions_every_variant(variant):
    layers = _conv_batch(10, 4321 + variant)
    ug.force_variant(ug.MODE_FP32, variant)
    try:
        plan, outs, keep = _conv_plan(layers, False)
    finally:
        ug.force_variant(ug.MODE_FP32, -1)
    plan.run(None)
    hip.synchronize()
    for (ich, h, w, k, pad, stride, ch, x, wts, bias, act), o in zip(layers, outs):
        exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias, act=act, slope=0.1)
        assert orc.check_bound(o.download(), exact, scale, orc.tolerance(ich * k * k))[0]
    plan.destroy()


@pytest.mark.parametrize("variant", [-1, 0, 1, 2, 3, 4])
def test_grouped_convolutions_fp32_mfma(variant):
    """Convolution plans on the FP32 matrix cores (the persistent kernel's
    FP32-MFMA tiles with the implicit im2col operand), captured; two replays
    bitwise identical, every layer within the plain FP32 bound."""
    mode = ug.MODE_FP32_MFMA
    layers = _conv_batch(16, 555 + variant)
    ug.force_variant(mode, variant)
    try:
        plan, outs, keep = _conv_plan(layers, True, mode=mode)
    finally:
        ug.force_variant(mode, -1)
    first = None
    for _ in range(2):
        plan.run(None)
        hip.synchronize()
        now = [o.download() for o in outs]
        if first is None:
            first = now
        else:
            assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(first, now))
    for (ich, h, w, k, pad, stride, ch, x, wts, bias, act), got in zip(layers, first):
        exact, scale = orc.conv_lrelu(x, ich, w, h, wts, k, pad, stride, ch, bias, act=act, slope=0.1)
        ok, worst = orc.check_bound(got, exact, scale, orc.tolerance(ich * k * k))
        assert ok, ((ich, h, w, k, pad, stride, ch), variant, worst)
    plan.destroy()


def test_grouped_convolution_errors():
    L = ug.load()
    h = ctypes.c_void_p()
    bad = ug.ConvProblem(16, 1, 4, 4, 16, 5, 0, 1, 16, 1, None, 1, 0.1)      # kernel larger than the image
    arr = (ug.ConvProblem * 1)(bad)
    assert L.ugemm_hip_conv_grouped_plan_create(ctypes.byref(h), 1, arr, 0, None) == ug.ERR_ARG
    bad = ug.ConvProblem(None, 1, 4, 4, 16, 3, 0, 1, 16, 1, None, 1, 0.1)    # NULL input
    arr = (ug.ConvProblem * 1)(bad)
    assert L.ugemm_hip_conv_grouped_plan_create(ctypes.byref(h), 1, arr, 0, None) == ug.ERR_ARG
    # Modes other than FP32 and FP32-MFMA are refused before anything is read.
    for mode in (ug.MODE_BF16X3, ug.MODE_FP8_BLOCK, 42):
        assert L.ugemm_hip_conv_grouped_plan_create_mode(ctypes.byref(h), 1, arr, mode, 0, None) == \
            ug.ERR_UNSUPPORTED
