"""CPU oracle for the GEMM path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module; the product library never does.

Contents
  * ctypes bindings to liboracle.so (oracle.c: the naive loops of
    ugemm.h:162-411, FP64 / double-double references with the error scale,
    the gemm_cpu.h blocking) and, when built, to oracle/_ref/libugemm_ref.so
    (the reference's own ugemm.h compiled by oracle/Makefile);
  * input generators for the workload distributions (numpy PCG64, seeded);
  * cmp_results: the error statistic of check_sgemm.c:56-85;
  * acceptance checks: |C - C_ref| <= tol * scale elementwise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libugemm_ref.so")

EPS32 = 2.0 ** -24   # unit roundoff of binary32
EPS64 = 2.0 ** -53

_o = None
_r = None


def _c(x):
    return x.encode() if isinstance(x, str) else x


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: run make -C oracle")
        L = ctypes.CDLL(ORACLE_SO)
        c, i, f, d, vp = ctypes.c_char, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
        for name, T in (("oracle_sgemm_naive", f), ("oracle_sgemm_packed", f)):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [c, c, c, i, i, i, T, vp, i, vp, i, T, vp, i]
        for name in ("oracle_dgemm_naive", "oracle_dgemm_packed"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [c, c, c, i, i, i, d, vp, i, vp, i, d, vp, i]
        for name in ("oracle_sgemm_f64", "oracle_dgemm_dd"):
            fn = getattr(L, name)
            fn.restype = i
            fn.argtypes = [c, c, c, i, i, i, d, vp, i, vp, i, d, vp, i, vp, i, vp, vp, i]
        L.oracle_saxpy.restype = None
        L.oracle_saxpy.argtypes = [i, f, vp, i, vp, i]
        L.oracle_daxpy.restype = None
        L.oracle_daxpy.argtypes = [i, d, vp, i, vp, i]
        L.oracle_saxpy3.restype = None
        L.oracle_saxpy3.argtypes = [i, f, vp, vp, vp]
        L.oracle_daxpy3.restype = None
        L.oracle_daxpy3.argtypes = [i, d, vp, vp, vp]
        L.oracle_sgemv.restype = None
        L.oracle_sgemv.argtypes = [c, i, i, f, vp, i, vp, i, f, vp, i, i]
        L.oracle_dgemv.restype = None
        L.oracle_dgemv.argtypes = [c, i, i, d, vp, i, vp, i, d, vp, i, i]
        L.oracle_im2col.restype = None
        L.oracle_im2col.argtypes = [vp, i, i, i, i, i, i, i, i, i, vp]
        L.oracle_conv_lrelu.restype = i
        L.oracle_conv_lrelu.argtypes = [vp, i, i, i, vp, i, i, i, vp, i, vp, i, d, vp]
        _o = L
    return _o


def ref_lib():
    """The reference's CPU library compiled from its sources (None if not built)."""
    global _r
    if _r is None:
        if not os.path.exists(REF_SO):
            return None
        L = ctypes.CDLL(REF_SO)
        c, i, f, d, vp = ctypes.c_char, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
        for name in ("sgemm_cpu", "sgemm_c", "sgemm_avx", "ref_sgemm_sse"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [c, c, c, i, i, i, f, vp, i, vp, i, f, vp, i]
        for name in ("dgemm_cpu", "_dgemm_c", "dgemm_avx"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [c, c, c, i, i, i, d, vp, i, vp, i, d, vp, i]
        L.saxpy_cpu.restype = None
        L.saxpy_cpu.argtypes = [i, f, vp, i, vp, i]
        L.daxpy_cpu.restype = None
        L.daxpy_cpu.argtypes = [i, d, vp, i, vp, i]
        L.saxpy_avx.restype = None
        L.saxpy_avx.argtypes = [i, f, vp, vp, vp]
        L.daxpy_avx.restype = None
        L.daxpy_avx.argtypes = [i, d, vp, vp, vp]
        L.sgemv_cpu.restype = None
        L.sgemv_cpu.argtypes = [c, i, i, f, vp, i, vp, i, f, vp, i]
        L.dgemv_cpu.restype = None
        L.dgemv_cpu.argtypes = [c, i, i, d, vp, i, vp, i, d, vp, i]
        _r = L
    return _r


def _p(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------
# Calls into oracle.c

def sgemm_naive(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    """sgemm_cpu semantics (ugemm.h:287-411); C is updated in place."""
    assert A.dtype == B.dtype == C.dtype == np.float32
    lib().oracle_sgemm_naive(_c(major), _c(ta), _c(tb), M, N, K, alpha, _p(A), lda, _p(B), ldb, beta, _p(C), ldc)


def dgemm_naive(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    assert A.dtype == B.dtype == C.dtype == np.float64
    lib().oracle_dgemm_naive(_c(major), _c(ta), _c(tb), M, N, K, alpha, _p(A), lda, _p(B), ldb, beta, _p(C), ldc)


def sgemm_packed(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    lib().oracle_sgemm_packed(_c(major), _c(ta), _c(tb), M, N, K, alpha, _p(A), lda, _p(B), ldb, beta, _p(C), ldc)


def dgemm_packed(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    lib().oracle_dgemm_packed(_c(major), _c(ta), _c(tb), M, N, K, alpha, _p(A), lda, _p(B), ldb, beta, _p(C), ldc)


def reference_f64(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C0, ldc, rows=None, nthreads=0):
    """Accurate result and error scale.

    float32 operands: compensated FP64 sums; float64 operands: double-double.
    Returns (out, scale) as float64 arrays of shape (len(rows) or M, N) for
    row-major, or (M, N) indexed [m, n] for column-major too (C[m, n]).
    `rows` selects rows m of C (row-major) -- used to sample large shapes.
    """
    dbl = A.dtype == np.float64
    fn = lib().oracle_dgemm_dd if dbl else lib().oracle_sgemm_f64
    r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nr = M if r is None else len(r)
    out = np.empty((nr, N), np.float64)
    scale = np.empty((nr, N), np.float64)
    c0 = C0 if (C0 is not None and beta != 0) else None
    rc = fn(_c(major), _c(ta), _c(tb), M, N, K, float(alpha), _p(A), lda, _p(B), ldb, float(beta), _p(c0), ldc,
            _p(r), nr, _p(out), _p(scale), nthreads)
    if rc != 0:
        raise ValueError("oracle: bad flags")
    return out, scale


# ---------------------------------------------------------------------------
# Inputs

def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def uniform(g, shape, lo=-1.0, hi=1.0, dtype=np.float32):
    """U[lo, hi) built from 24 random bits: exact in binary32 for [-1, 1)."""
    k = g.integers(0, 1 << 24, size=shape, dtype=np.int64).astype(np.float64)
    return (lo + (hi - lo) * k * 2.0 ** -24).astype(dtype)


def normal(g, shape, dtype=np.float32):
    return g.standard_normal(size=shape).astype(dtype)


def wide_range(g, shape, emin=-12, emax=12, dtype=np.float32):
    """sign * 2^e * (1 + u), e ~ U{emin..emax}, u ~ U[0,1) on 23 bits (exact)."""
    s = np.where(g.integers(0, 2, size=shape) == 0, -1.0, 1.0)
    e = g.integers(emin, emax + 1, size=shape)
    u = g.integers(0, 1 << 23, size=shape).astype(np.float64) * 2.0 ** -23
    return (s * np.ldexp(1.0 + u, e)).astype(dtype)


def stored_shape(major, trans, rows, cols, ld):
    """Shape of the buffer holding op(X) (rows x cols) with leading dimension ld.

    Row-major 'N' stores rows x ld; 'T' stores cols x ld.  Column-major is
    the transpose of that picture (buffer of  cols x ld for 'N')."""
    t = trans != "N"
    if major == "R":
        return (cols, ld) if t else (rows, ld)
    return (rows, ld) if t else (cols, ld)


def op_matrix(major, trans, X, rows, cols):
    """The logical rows x cols matrix op(X) seen through a stored buffer X."""
    t = trans != "N"
    if major == "R":
        return X[:cols, :rows].T if t else X[:rows, :cols]
    return X[:rows, :cols] if t else X[:cols, :rows].T


def c_matrix(major, C, M, N):
    return C[:M, :N] if major == "R" else C[:N, :M].T


# ---------------------------------------------------------------------------
# Statistics and acceptance

def cmp_results(M, N, ref, res, ld):
    """check_sgemm.c:56-85: (stdErr, stdRef, maxErr, argmax (m, n), verdict)."""
    r = ref.reshape(-1, ld)[:M, :N].astype(np.float64)
    x = res.reshape(-1, ld)[:M, :N].astype(np.float64)
    err = x - r
    a = np.abs(err)
    i = int(np.argmax(a))
    max_err = float(a.flat[i])
    std_err = float(np.sqrt((err * err).sum() / (M * N)))
    s1, s2 = float(r.sum()), float((r * r).sum())
    std_ref = float(np.sqrt(max(s2 * (M * N) - s1 * s1, 0.0)) / (M * N))
    verdict = "FAIL !!!" if max_err > std_ref * 1e-5 else (
        "Sucks !" if (max_err > std_ref * 3e-5 or std_err > std_ref * 1e-6) else "")
    return std_err, std_ref, max_err, divmod(i, N), verdict


def tolerance(K, eps=EPS32, extra=0.0):
    """Elementwise acceptance factor: |C - C_exact| <= tolerance * scale.

    2*K*eps covers sequential accumulation of K products (the bound used for
    the plain path); 2*eps more covers the alpha/beta epilogue roundings;
    `extra` adds a per-product term for reduced-precision operand forms."""
    return 2.0 * (K + 1) * eps + extra


def check_bound(res, exact, scale, tol):
    """Returns (ok, worst ratio |res-exact| / (tol*scale))."""
    res = np.asarray(res, np.float64)
    err = np.abs(res - exact)
    lim = tol * scale
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(lim > 0, err / lim, np.where(err > 0, np.inf, 0.0))
    worst = float(np.nanmax(ratio)) if ratio.size else 0.0
    ok = bool(np.all(np.isfinite(res) == np.isfinite(exact))) and worst <= 1.0
    return ok, worst


def rel_errors(res, exact):
    """max and mean of |res - exact| / |exact| over entries with exact != 0."""
    res = np.asarray(res, np.float64)
    nz = exact != 0
    rel = np.abs(res[nz] - exact[nz]) / np.abs(exact[nz])
    return (float(rel.max()), float(rel.mean())) if rel.size else (0.0, 0.0)


# ---------------------------------------------------------------------------
# FP8 (e4m3, OCP "E4M3FN": bias 7, no infinities, max 448, subnormal step
# 2^-9) with one power-of-two scale per 128 x 128 block.  The reference has
# no FP8 path; this restates the build's own quantisation rule so that the
# FP8 mode can be checked element by element.  Pinned by the e4m3 definition
# (tests/test_oracle.py: known encodings, rounding ties, subnormals).

def e4m3_table():
    """All finite non-negative e4m3 values in code order (code = 8*E + M)."""
    vals = []
    for e in range(16):
        for m in range(8):
            if e == 15 and m == 7:
                continue  # NaN encoding
            vals.append(m * 2.0 ** -9 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7))
    return np.array(vals)


_E4M3 = None


def quantize_e4m3(x):
    """Round-to-nearest-even onto e4m3 (|x| <= 448 assumed, larger saturates)."""
    global _E4M3
    if _E4M3 is None:
        _E4M3 = e4m3_table()
    v = _E4M3
    x = np.asarray(x, np.float64)
    a = np.minimum(np.abs(x), v[-1])
    hi = np.clip(np.searchsorted(v, a, side="left"), 0, len(v) - 1)
    lo = np.clip(hi - 1, 0, len(v) - 1)
    d_lo, d_hi = a - v[lo], v[hi] - a
    pick_hi = (d_hi < d_lo) | ((d_hi == d_lo) & (hi % 2 == 0))
    q = np.where(pick_hi, v[hi], v[lo])
    return np.sign(x) * q


def block_exponent(amax):
    """Smallest e with amax / 2^e <= 448 (0 for an all-zero block)."""
    if not amax > 0:
        return 0
    m, e = np.frexp(amax / 448.0)
    if m == 0.5:
        e -= 1
    return int(min(max(e, -127), 127))


def fp8_block_dequant(X, block=128):
    """X (rows x K, logical) -> the values the FP8 mode multiplies: each
    block x block tile quantised with its power-of-two scale, then rescaled."""
    X = np.asarray(X, np.float64)
    out = np.empty_like(X)
    for r0 in range(0, X.shape[0], block):
        for k0 in range(0, X.shape[1], block):
            t = X[r0:r0 + block, k0:k0 + block]
            e = block_exponent(float(np.max(np.abs(t))) if t.size else 0.0)
            out[r0:r0 + block, k0:k0 + block] = quantize_e4m3(t / 2.0 ** e) * 2.0 ** e
    return out


# ---------------------------------------------------------------------------
# FP16X3: two-way fp16 split with a power-of-two scale per row of op(A) and
# per column of op(B).  The reference has no such path; this restates the
# build's own rule so that the form can be checked element by element.
# Pinned (tests/test_oracle.py) by Python's struct IEEE-half codec, exact
# small-integer inputs and the split's reconstruction bound.

def fp16_row_exponent(amax):
    """e such that a row with largest magnitude amax, scaled by 2^-e, lies in
    (-2^15, 2^15); -15 for 0, inf and NaN."""
    amax = float(amax)
    if amax == 0.0 or not np.isfinite(amax):
        return -15
    return int(np.frexp(amax)[1]) - 15


def fp16_split_rows(X):
    """X (rows x K, logical, float32 values) -> (hi, lo, e): float64 arrays of
    the fp16 planes of X_r * 2^-e_r (round to nearest even, subnormals kept)
    and the per-row exponents."""
    X = np.asarray(X, np.float32).astype(np.float64)
    e = np.array([fp16_row_exponent(np.max(np.abs(r))) if r.size else -15 for r in X], np.int64)
    Xs = np.ldexp(X, -e[:, None]).astype(np.float32)          # exact: only the exponent changes
    with np.errstate(over="ignore", invalid="ignore"):
        hi = Xs.astype(np.float16)
        r = Xs - hi.astype(np.float32)                          # exact in float32
        r = np.where(np.isfinite(hi), r, np.float32(0.0))
        lo = r.astype(np.float16)
    return hi.astype(np.float64), lo.astype(np.float64), e


def fp16x3_product(Aop, Bop):
    """The value FP16X3 computes for op(A) @ op(B) before alpha/beta, in
    float64: 2^(e_m + e_n) * (Ahi Bhi + Ahi Blo + Alo Bhi)."""
    ah, al, ea = fp16_split_rows(Aop)
    bh, bl, eb = fp16_split_rows(np.asarray(Bop).T)
    P = ah @ bh.T + ah @ bl.T + al @ bh.T
    return np.ldexp(P, ea[:, None] + eb[None, :])


# ---------------------------------------------------------------------------
# BF16X1 / BF16X3 / BF16X6: the NP-plane bf16 split of every FP32 operand and
# the product set the split core issues.  The reference has no such path;
# this restates the build's own rule so that the forms can be checked element
# by element.  Pinned (tests/test_oracle.py) by an independent nearest-value
# search for the rounding, the planes' reconstruction bounds, exact products
# of small integers, and a planted truncating split that must fail.

def bf16_round(x):
    """float32 values -> the nearest bf16 value (8 significant bits, FP32's
    exponent range), ties to even, computed on the float32 bit pattern;
    returned as float32.  Infinities and NaN pass unchanged; a finite value
    that rounds past the largest bf16 becomes infinite."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return np.where(np.isfinite(x), r.view(np.float32), x)


def bf16_split(X, NP):
    """X (float32 values) -> NP float64 planes h0 = bf16(x), h1 = bf16(x - h0),
    h2 = bf16(x - h0 - h1); the differences are exact in float32.  The low
    planes of a non-finite x are 0."""
    X = np.asarray(X, np.float32)
    h = bf16_round(X)
    planes = [h.astype(np.float64)]
    if NP > 1:
        with np.errstate(invalid="ignore"):
            r = np.where(np.isfinite(h), X - h, np.float32(0.0)).astype(np.float32)
        l1 = bf16_round(r)
        planes.append(l1.astype(np.float64))
        if NP > 2:
            planes.append(bf16_round((r - l1).astype(np.float32)).astype(np.float64))
    return planes


def bf16_products(NP):
    """The (i, j) plane pairs the split core multiplies: i + j <= NP - 1
    (NP = 1: h0h0; 2: + h0h1, h1h0; 3: + h1h1, h0h2, h2h0)."""
    return [(i, j) for i in range(NP) for j in range(NP) if i + j <= NP - 1]


def bf16_product(Aop, Bop, NP):
    """The value a BF16 form with NP planes computes for op(A) @ op(B) before
    alpha/beta, in float64: sum over bf16_products(NP) of A_i @ B_j."""
    ap = bf16_split(Aop, NP)
    bp = bf16_split(np.asarray(Bop, np.float32).T, NP)
    P = None
    for (i, j) in bf16_products(NP):
        t = ap[i] @ bp[j].T
        P = t if P is None else P + t
    return P


# ---------------------------------------------------------------------------
# Convolution (sgemm_gl1.h:166-218): bindings to oracle.c.

def conv_out_shape(h, w, k, pad, stride):
    return (h + 2 * pad - k) // stride + 1, (w + 2 * pad - k) // stride + 1


def im2col(im, ich, h, w, k, pad, stride):
    """(ich*k*k, hcol*wcol) column matrix of a CHW float32 image."""
    hcol, wcol = conv_out_shape(h, w, k, pad, stride)
    im = np.ascontiguousarray(im, np.float32)
    col = np.zeros((ich * k * k, hcol * wcol), np.float32)
    lib().oracle_im2col(_p(im), ich, h, w, k, k, pad, pad, stride, stride, _p(col))
    return col


def conv_lrelu(inputs, ich, w, h, weights, k, pad, stride, ch, bias=None, act=1, slope=0.1):
    """(out, scale): float64 [ch, hcol*wcol] of gl_convolution_LReLU's result
    and its elementwise error scale."""
    hcol, wcol = conv_out_shape(h, w, k, pad, stride)
    inputs = np.ascontiguousarray(inputs, np.float32)
    weights = np.ascontiguousarray(weights, np.float32)
    b = None if bias is None else np.ascontiguousarray(bias, np.float32)
    out = np.zeros((ch, hcol * wcol), np.float64)
    scale = np.zeros((ch, hcol * wcol), np.float64)
    rc = lib().oracle_conv_lrelu(_p(inputs), ich, w, h, _p(weights), k, pad, stride, _p(out), ch,
                                 None if b is None else _p(b), act, slope, _p(scale))
    if rc:
        raise MemoryError("oracle_conv_lrelu")
    return out, scale


# ---------------------------------------------------------------------------
# BLAS-1/2 (ugemm.h:76-160): bindings to oracle.c and dense views.

def axpy3(N, a, x, y):
    """z = a*x + y with one rounding per element (oracle.c: the FMA branch of
    saxpy_avx / daxpy_avx, ugemm.h:41-74); float32 or float64 by x's dtype."""
    dt = np.float64 if x.dtype == np.float64 else np.float32
    x, y = np.ascontiguousarray(x, dt), np.ascontiguousarray(y, dt)
    z = np.empty(N, dt)
    (lib().oracle_daxpy3 if dt == np.float64 else lib().oracle_saxpy3)(N, a, _p(x), _p(y), _p(z))
    return z


def axpy(N, alpha, x, incx, y, incy):
    """y updated in place (working precision of y's dtype)."""
    fn = lib().oracle_daxpy if y.dtype == np.float64 else lib().oracle_saxpy
    fn(N, float(alpha), _p(x), incx, _p(y), incy)


def gemv(trans, M, N, alpha, A, lda, x, incx, beta, y, incy, x_by_incy=False):
    """y updated in place; x_by_incy reproduces the reference's x indexing."""
    fn = lib().oracle_dgemv if y.dtype == np.float64 else lib().oracle_sgemv
    fn(_c(trans), M, N, float(alpha), _p(A), lda, _p(x), incx, float(beta), _p(y), incy, 1 if x_by_incy else 0)


def gemv_dense(trans, M, N, A, lda, x, incx):
    """(a, xv): the M x N matrix a(m, n) and the N-vector x[n*incx] that the
    gemv of ugemm.h:103-160 multiplies (for an accurate reference)."""
    A = np.asarray(A).ravel()
    m, n = np.arange(M)[:, None], np.arange(N)[None, :]
    idx = (m + n * lda) if trans == "N" else (n + m * lda)
    return A[idx], np.asarray(x).ravel()[np.arange(N) * incx]
