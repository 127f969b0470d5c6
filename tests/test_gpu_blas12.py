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
