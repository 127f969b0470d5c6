"""BLAS-1/2 helpers (saxpy/daxpy/sgemv/dgemv, ugemm.h:76-160) on the GPU vs
the oracle and the reference's golden vectors.

Acceptance: axpy within 2 eps (|y0| + |alpha x|) elementwise (one fused or
two separate roundings); gemv within 2(N+2) eps (|alpha| |A||x| + |beta||y0|)
against an FP64 (FP32 inputs) or double-double (FP64 inputs) evaluation.
The golden vectors are compared only where incx == incy: elsewhere the
reference reads x with incy (ugemm.h:116,147), which this build corrects."""
import ast
import os

import numpy as np
import pytest

import ugemm_amd as ug
from ugemm_amd import hip
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _accurate_gemv(trans, M, N, alpha, A, lda, x, incx, beta, y0, incy):
    a, xv = orc.gemv_dense(trans, M, N, A, lda, x, incx)
    y0v = y0[::incy][:M].copy()
    if N == 0:
        return beta * y0v.astype(np.float64), np.abs(beta * y0v.astype(np.float64))
    a = np.ascontiguousarray(a)
    xm = np.ascontiguousarray(xv.reshape(N, 1))
    c0 = np.ascontiguousarray(y0v.reshape(M, 1))
    out, scale = orc.reference_f64("R", "N", "N", M, 1, N, alpha, a, N, xm, 1, beta, c0, 1)
    return out[:, 0], scale[:, 0]


def _gemv_case(dt, trans, M, N, pad, incx, incy, alpha, beta, seed):
    g = orc.rng(seed)
    rows, cols = (M, N) if trans == "N" else (N, M)
    lda = max(1, rows + pad)
    A = orc.uniform(g, (max(1, cols) * lda,), dtype=dt)
    x = orc.uniform(g, (max(1, (N - 1) * incx + 1),), dtype=dt)
    y0 = orc.uniform(g, ((M - 1) * incy + 1,), dtype=dt)
    return A, lda, x, y0


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("M,N", [(1, 1), (7, 3), (300, 257), (1000, 17), (17, 3000), (4096, 2048)])
@pytest.mark.parametrize("incx,incy", [(1, 1), (3, 2)])
def test_gemv_device(dt, trans, M, N, incx, incy):
    alpha, beta = 1.5, 0.5
    A, lda, x, y0 = _gemv_case(dt, trans, M, N, 3, incx, incy, alpha, beta, M * 7 + N + incx)
    dA, dx, dy = hip.DeviceArray.from_host(A), hip.DeviceArray.from_host(x), hip.DeviceArray.from_host(y0)
    print("gemv", dt.__name__, trans, M, N, incx, incy, flush=True)
    ug.gemv_dev(trans, M, N, alpha, dA, lda, dx, incx, beta, dy, incy, double=dt == np.float64)
    hip.synchronize()
    y = dy.download()
    want, scale = _accurate_gemv(trans, M, N, alpha, A, lda, x, incx, beta, y0, incy)
    eps = orc.EPS64 if dt == np.float64 else orc.EPS32
    err = np.abs(y[::incy][:M].astype(np.float64) - want)
    assert np.all(err <= 2 * (N + 2) * eps * scale), float(np.max(err / np.maximum(scale, 1e-300)))
    if incy > 1:
        assert np.array_equal(y[1::incy], y0[1::incy])     # gaps of a strided y untouched


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("N", [1, 5, 1000, 1 << 20, (1 << 20) + 3])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, 3)])
def test_axpy_device(dt, N, incx, incy):
    g = orc.rng(N + incx)
    x = orc.uniform(g, ((N - 1) * incx + 1,), dtype=dt)
    y0 = orc.uniform(g, ((N - 1) * incy + 1,), dtype=dt)
    dx, dy = hip.DeviceArray.from_host(x), hip.DeviceArray.from_host(y0)
    ug.axpy_dev(N, -0.75, dx, incx, dy, incy, double=dt == np.float64)
    hip.synchronize()
    y = dy.download().astype(np.float64)
    want = y0.astype(np.float64).copy()
    want[::incy][:N] += -0.75 * x[::incx][:N].astype(np.float64)
    sc = np.abs(y0.astype(np.float64))
    sc[::incy][:N] += 0.75 * np.abs(x[::incx][:N].astype(np.float64))
    eps = orc.EPS64 if dt == np.float64 else orc.EPS32
    assert np.all(np.abs(y - want) <= 2 * eps * sc)


def test_host_entries_and_golden_vectors():
    """saxpy_hip/daxpy_hip/sgemv_hip/dgemv_hip (host pointers, the
    reference's argument lists) against the reference's own outputs."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_blas12.npz"))
    n = 0
    for key, name, N, incx, incy, alpha, beta, trans, M, lda, xbug in (ast.literal_eval(s) for s in d["index"]):
        if xbug:
            continue
        dt = np.float64 if name[0] == "d" else np.float32
        eps = orc.EPS64 if dt == np.float64 else orc.EPS32
        x, y0, yref = d[key + "_x"], d[key + "_y0"], d[key + "_y"]
        y = y0.copy()
        if "axpy" in name:
            ug.axpy(N, alpha, x, incx, y, incy)
            sc = np.abs(y0.astype(np.float64))
            sc[::incy][:N] += abs(alpha) * np.abs(x[::incx][:N].astype(np.float64))
            assert np.all(np.abs(y.astype(np.float64) - yref) <= 4 * eps * sc), key
        else:
            A = d[key + "_A"]
            ug.gemv(trans, M, N, alpha, A, lda, x, incx, beta, y, incy)
            _, scale = _accurate_gemv(trans, M, N, alpha, A, lda, x, incx, beta, y0, incy)
            err = np.abs(y[::incy][:M].astype(np.float64) - yref[::incy][:M])
            assert np.all(err <= 4 * (N + 2) * eps * scale), key
        n += 1
    assert n > 100


def test_gemv_beta_zero_never_reads_y_and_alpha_zero():
    M, N = 50, 40
    A, lda, x, y0 = _gemv_case(np.float32, "N", M, N, 0, 1, 1, 1.0, 0.0, 5)
    y = np.full(M, np.nan, np.float32)
    ug.gemv("N", M, N, 1.0, A, lda, x, 1, 0.0, y, 1)
    assert np.all(np.isfinite(y))
    y = y0.copy()
    ug.gemv("T", M, N, 0.0, A, N, x, 1, 2.0, y, 1)
    assert np.array_equal(y, 2.0 * y0)


def test_blas12_argument_errors():
    L = ug.load()
    d = hip.DeviceArray(64, np.float32)
    assert L.ugemm_hip_saxpy_dev(4, 1.0, d.ptr, 0, d.ptr, 1, None) == ug.ERR_ARG
    assert L.ugemm_hip_sgemv_dev(b"N", 8, 4, 1.0, d.ptr, 4, d.ptr, 1, 0.0, d.ptr, 1, None) == ug.ERR_ARG   # lda < M
    assert L.ugemm_hip_sgemv_dev(b"T", 4, 8, 1.0, d.ptr, 8, d.ptr, -1, 0.0, d.ptr, 1, None) == ug.ERR_ARG
    assert L.ugemm_hip_saxpy_dev(0, 1.0, None, 1, None, 1, None) == ug.OK


# ---------------------------------------------------------------------------
# Three-operand axpy (saxpy_avx / daxpy_avx, ugemm.h:41-74): z = a*x + y with
# one rounding, compared bit for bit (both sides are correctly rounded FMAs).

def _axpy3_golden():
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_axpy3.npz"))
    return d, [ast.literal_eval(t) for t in d["index"]]


def test_axpy3_host_entry_matches_reference_golden():
    """saxpy_avx_hip / daxpy_avx_hip on the golden inputs: the first N outputs
    equal the reference's bit for bit, and nothing beyond N is written (the
    reference's whole-vector loop writes up to the next multiple of 8 / 4)."""
    d, index = _axpy3_golden()
    for key, name, N, a in index:
        x, y, zfull = d[key + "_x"], d[key + "_y"], d[key + "_zfull"]
        z = np.full(len(zfull), 7.0, x.dtype)
        ug.axpy3(N, a, x, y, z)
        assert np.array_equal(z[:N].view(np.uint8), zfull[:N].view(np.uint8)), key
        assert np.all(z[N:] == 7.0), key


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("N,offset", [(1, 0), (5, 1), (1000, 0), (1000, 3), ((1 << 22) + 3, 0), ((1 << 20) + 7, 1)])
def test_axpy3_device_entry_vs_oracle(dt, N, offset):
    """Device entry, aligned (vector kernel) and misaligned (scalar kernel)
    starts, in place (z = y) and out of place, vs the oracle's FMA loop."""
    g = orc.rng(N + offset + (7 if dt == np.float64 else 0))
    x, y = orc.uniform(g, (N + offset,), dtype=dt), orc.uniform(g, (N + offset,), dtype=dt)
    want = orc.axpy3(N, -0.7, x[offset:], y[offset:])
    dx, dy, dz = hip.DeviceArray.from_host(x), hip.DeviceArray.from_host(y), hip.DeviceArray((N + offset,), dt)
    try:
        ug.axpy3_dev(N, -0.7, dx.at(offset), dy.at(offset), dz.at(offset), double=dt == np.float64)
        hip.synchronize()
        assert np.array_equal(dz.download()[offset:], want)
        ug.axpy3_dev(N, -0.7, dx.at(offset), dy.at(offset), dy.at(offset), double=dt == np.float64)   # z = y
        hip.synchronize()
        assert np.array_equal(dy.download()[offset:], want)
    finally:
        for d in (dx, dy, dz):
            d.free()


def test_axpy3_argument_errors():
    L = ug.load()
    x = np.zeros(4, np.float32)
    L.saxpy_avx_hip(-1, 1.0, x.ctypes.data, x.ctypes.data, x.ctypes.data)
    assert L.ugemm_hip_last_error() == ug.ERR_ARG
    L.saxpy_avx_hip(4, 1.0, x.ctypes.data, None, x.ctypes.data)
    assert L.ugemm_hip_last_error() == ug.ERR_ARG
    L.saxpy_avx_hip(0, 1.0, None, None, None)
    assert L.ugemm_hip_last_error() == ug.OK
