"""ugemm_amd: Python access to libugemm_hip.so, the MI355X (gfx950) GEMM
backend for the ugemm C API.

The library is the product; this module only loads it with ctypes and gives
Python callers the same argument lists as the C ABI (include/ugemm_hip.h):

    sgemm(major, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc)

mirrors the uut signature of sgemm_cpu / sgemm_avx (ugemm.h:287-301) on host
numpy arrays, and sgemm_dev(...) the device-pointer, stream-ordered entry.
Errors raise UgemmError carrying the library's status code and message.
There is no CPU fallback: if libugemm_hip.so is missing, importing the
bindings (load()) raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libugemm_hip.so")

MODE_FP32, MODE_BF16X3, MODE_BF16X1, MODE_FP8_BLOCK, MODE_BF16X6, MODE_FP16X3, MODE_FP32_MFMA = 0, 1, 2, 3, 4, 5, 6
MODE_DGEMM = -1  # variant tables of the DGEMM family
MODES = {"fp32": MODE_FP32, "bf16x3": MODE_BF16X3, "bf16x1": MODE_BF16X1,
         "fp8": MODE_FP8_BLOCK, "bf16x6": MODE_BF16X6, "fp16x3": MODE_FP16X3, "fp32_mfma": MODE_FP32_MFMA}

OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_NODEV, ERR_UNSUPPORTED, ERR_TUNE_FILE, ERR_SCRATCH = 0, -1, -2, -3, -4, -5, -6, -7
# Size of the static device scratch of ugemm_hip_sgemm_noalloc_dev (UGEMM_HIP_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 64 << 20

_lib = None


class UgemmError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"ugemm_hip status {status}: {message}")
        self.status = status


class Problem(ctypes.Structure):
    """ugemm_problem_t (include/ugemm_hip.h)."""
    _fields_ = [("transa", ctypes.c_char), ("transb", ctypes.c_char),
                ("M", ctypes.c_int), ("N", ctypes.c_int), ("K", ctypes.c_int),
                ("alpha", ctypes.c_float), ("beta", ctypes.c_float),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_int)]


class ConvProblem(ctypes.Structure):
    """ugemm_conv_problem_t (include/ugemm_hip.h)."""
    _fields_ = [("in_", ctypes.c_void_p), ("ich", ctypes.c_int), ("w", ctypes.c_int), ("h", ctypes.c_int),
                ("weights", ctypes.c_void_p), ("k", ctypes.c_int), ("pad", ctypes.c_int), ("stride", ctypes.c_int),
                ("out", ctypes.c_void_p), ("ch", ctypes.c_int),
                ("bias", ctypes.c_void_p), ("act", ctypes.c_int), ("slope", ctypes.c_float)]


def load(path=None):
    """Loads libugemm_hip.so (once) and declares every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(path)
    c, i, f, d, vp, sz = ctypes.c_char, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
    cs = ctypes.c_char_p
    sig = {
        "sgemm_hip": (None, [c, c, c, i, i, i, f, vp, i, vp, i, f, vp, i]),
        "dgemm_hip": (None, [c, c, c, i, i, i, d, vp, i, vp, i, d, vp, i]),
        "ugemm_hip_sgemm": (i, [c, c, c, i, i, i, f, vp, i, vp, i, f, vp, i, i]),
        "ugemm_hip_init": (i, [i, sz]),
        "ugemm_hip_finish": (None, []),
        "sgemm_hip_packed": (None, [c, c, i, i, i, f, vp, vp, f, vp]),
        "ugemm_hip_sgemm_dev": (i, [c, c, c, i, i, i, f, vp, i, vp, i, f, vp, i, i, vp]),
        "ugemm_hip_dgemm_dev": (i, [c, c, c, i, i, i, d, vp, i, vp, i, d, vp, i, vp]),
        "ugemm_hip_sgemm_noalloc_dev": (i, [c, c, c, i, i, i, f, vp, i, vp, i, f, vp, i, i, vp]),
        "ugemm_hip_sgemm_ex_dev": (i, [c, c, c, i, i, i, f, vp, i, vp, i, f, vp, i, i, vp, i, i, f, vp]),
        "ugemm_hip_conv2d_dev": (i, [vp, i, i, i, vp, i, i, i, vp, i, vp, i, f, i, vp]),
        "convolution_LReLU_hip": (None, [vp, i, i, i, vp, i, i, i, vp, i, vp]),
        "ugemm_hip_grouped_plan_create": (i, [ctypes.POINTER(vp), i, ctypes.POINTER(Problem), i, i, vp]),
        "ugemm_hip_grouped_run": (i, [vp, vp]),
        "ugemm_hip_conv_grouped_plan_create": (i, [ctypes.POINTER(vp), i, ctypes.POINTER(ConvProblem), i, vp]),
        "ugemm_hip_conv_grouped_plan_create_mode": (i, [ctypes.POINTER(vp), i, ctypes.POINTER(ConvProblem), i, i, vp]),
        "ugemm_hip_grouped_plan_destroy": (None, [vp]),
        "ugemm_hip_sgemm_grouped": (i, [i, ctypes.POINTER(Problem), i, vp]),
        "ugemm_hip_last_error": (i, []),
        "ugemm_hip_last_error_message": (cs, []),
        "ugemm_hip_status_string": (cs, [i]),
        "ugemm_hip_tune_sgemm": (i, [c, c, i, i, i, i]),
        "ugemm_hip_tune_grouped": (i, [vp]),
        "ugemm_hip_grouped_variant_count": (i, []),
        "ugemm_hip_grouped_variant_name": (cs, [i]),
        "ugemm_hip_grouped_mode_variant_count": (i, [i]),
        "ugemm_hip_grouped_mode_variant_name": (cs, [i, i]),
        "ugemm_hip_tune_load": (i, [cs]),
        "ugemm_hip_tune_save": (i, [cs]),
        "ugemm_hip_tune_lookup": (i, [c, c, i, i, i, i]),
        "ugemm_hip_force_variant": (i, [i, i]),
        "ugemm_hip_force_splits": (i, [i, i]),
        "ugemm_hip_variant_count": (i, [i]),
        "ugemm_hip_variant_name": (cs, [i, i]),
        "ugemm_hip_version": (cs, []),
        "saxpy_hip": (None, [i, f, vp, i, vp, i]),
        "daxpy_hip": (None, [i, d, vp, i, vp, i]),
        "saxpy_avx_hip": (None, [i, f, vp, vp, vp]),
        "daxpy_avx_hip": (None, [i, d, vp, vp, vp]),
        "ugemm_hip_saxpy3_dev": (i, [i, f, vp, vp, vp, vp]),
        "ugemm_hip_daxpy3_dev": (i, [i, d, vp, vp, vp, vp]),
        "sgemv_hip": (None, [c, i, i, f, vp, i, vp, i, f, vp, i]),
        "dgemv_hip": (None, [c, i, i, d, vp, i, vp, i, d, vp, i]),
        "ugemm_hip_saxpy_dev": (i, [i, f, vp, i, vp, i, vp]),
        "ugemm_hip_daxpy_dev": (i, [i, d, vp, i, vp, i, vp]),
        "ugemm_hip_sgemv_dev": (i, [c, i, i, f, vp, i, vp, i, f, vp, i, vp]),
        "ugemm_hip_dgemv_dev": (i, [c, i, i, d, vp, i, vp, i, d, vp, i, vp]),
        "ugemm_hip_timing_enable": (i, [i]),
        "ugemm_hip_timing_collect": (i, [ctypes.POINTER(ctypes.c_double)]),
        "ugemm_hip_mfma_ceiling": (i, [ctypes.c_void_p, i, i, ctypes.POINTER(ctypes.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _b(ch):
    return ch.encode() if isinstance(ch, str) else ch


def _raise_if(rc):
    if rc != OK:
        L = load()
        raise UgemmError(rc, L.ugemm_hip_last_error_message().decode())
    return rc


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "ptr"):
        return x.ptr
    raise TypeError(f"cannot take the address of {type(x)}")


def init(device=-1, max_bytes=0):
    return _raise_if(load().ugemm_hip_init(device, max_bytes))


def finish():
    load().ugemm_hip_finish()


def sgemm(major, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode=MODE_FP32):
    """Host-pointer SGEMM (synchronous); C (float32 numpy array) is updated in place."""
    L = load()
    rc = L.ugemm_hip_sgemm(_b(major), _b(transa), _b(transb), M, N, K, alpha, _ptr(A), lda, _ptr(B), ldb,
                           beta, _ptr(C), ldc, mode)
    return _raise_if(rc)


def dgemm(major, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    """Host-pointer DGEMM through the void uut entry dgemm_hip; raises on error."""
    L = load()
    L.dgemm_hip(_b(major), _b(transa), _b(transb), M, N, K, alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(C), ldc)
    return _raise_if(L.ugemm_hip_last_error())


def sgemm_dev(major, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, mode=MODE_FP32, stream=None):
    """Device-pointer SGEMM on `stream` (a ugemm_amd.hip.Stream, a raw handle, or None)."""
    s = stream.handle if hasattr(stream, "handle") else stream
    rc = load().ugemm_hip_sgemm_dev(_b(major), _b(transa), _b(transb), M, N, K, alpha, _ptr(dA), lda,
                                    _ptr(dB), ldb, beta, _ptr(dC), ldc, mode, s)
    return _raise_if(rc)


def sgemm_noalloc_dev(major, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, mode=MODE_FP32,
                      stream=None):
    """sgemm_dev whose workspace comes from the library's static device scratch
    (no device allocation); raises UgemmError(ERR_SCRATCH) for a call whose
    workspace exceeds NOALLOC_SCRATCH_BYTES."""
    s = stream.handle if hasattr(stream, "handle") else stream
    rc = load().ugemm_hip_sgemm_noalloc_dev(_b(major), _b(transa), _b(transb), M, N, K, alpha, _ptr(dA), lda,
                                            _ptr(dB), ldb, beta, _ptr(dC), ldc, mode, s)
    return _raise_if(rc)


def dgemm_dev(major, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, stream=None):
    s = stream.handle if hasattr(stream, "handle") else stream
    rc = load().ugemm_hip_dgemm_dev(_b(major), _b(transa), _b(transb), M, N, K, alpha, _ptr(dA), lda,
                                    _ptr(dB), ldb, beta, _ptr(dC), ldc, s)
    return _raise_if(rc)


BIAS_ROWS, BIAS_COLS = 0, 1
ACT_NONE, ACT_LEAKY_RELU = 0, 1


def sgemm_ex_dev(major, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, mode=MODE_FP32,
                 bias=None, bias_dim=BIAS_ROWS, act=ACT_NONE, slope=0.1, stream=None):
    """sgemm_dev followed by the fused bias + activation epilogue."""
    s = stream.handle if hasattr(stream, "handle") else stream
    rc = load().ugemm_hip_sgemm_ex_dev(_b(major), _b(transa), _b(transb), M, N, K, alpha, _ptr(dA), lda, _ptr(dB),
                                       ldb, beta, _ptr(dC), ldc, mode, _ptr(bias), bias_dim, act, slope, s)
    return _raise_if(rc)


def conv2d_dev(d_in, ich, w, h, d_weights, k, pad, stride, d_out, ch, bias=None, act=ACT_LEAKY_RELU, slope=0.1,
               mode=MODE_FP32, stream=None):
    """Implicit-GEMM convolution + bias + activation on device buffers."""
    s = stream.handle if hasattr(stream, "handle") else stream
    rc = load().ugemm_hip_conv2d_dev(_ptr(d_in), ich, w, h, _ptr(d_weights), k, pad, stride, _ptr(d_out), ch,
                                     _ptr(bias), act, slope, mode, s)
    return _raise_if(rc)


def convolution_LReLU(inputs, ich, w, h, weights, k, pad, stride, outputs, ch, bias):
    """Host-array convolution with the argument list of gl_convolution_LReLU."""
    L = load()
    L.convolution_LReLU_hip(_ptr(inputs), ich, w, h, _ptr(weights), k, pad, stride, _ptr(outputs), ch, _ptr(bias))
    return _raise_if(L.ugemm_hip_last_error())


def axpy(N, alpha, x, incx, y, incy):
    """Host-array saxpy_hip / daxpy_hip (by y's dtype); y updated in place."""
    L = load()
    fn = L.daxpy_hip if y.dtype == np.float64 else L.saxpy_hip
    fn(N, alpha, _ptr(x), incx, _ptr(y), incy)
    return _raise_if(L.ugemm_hip_last_error())


def axpy3(N, a, x, y, z):
    """Host-array saxpy_avx_hip / daxpy_avx_hip (by z's dtype): z = a*x + y."""
    L = load()
    fn = L.daxpy_avx_hip if z.dtype == np.float64 else L.saxpy_avx_hip
    fn(N, a, _ptr(x), _ptr(y), _ptr(z))
    return _raise_if(L.ugemm_hip_last_error())


def axpy3_dev(N, a, dx, dy, dz, double=False, stream=None):
    s = stream.handle if hasattr(stream, "handle") else stream
    L = load()
    fn = L.ugemm_hip_daxpy3_dev if double else L.ugemm_hip_saxpy3_dev
    return _raise_if(fn(N, a, _ptr(dx), _ptr(dy), _ptr(dz), s))


def gemv(trans, M, N, alpha, A, lda, x, incx, beta, y, incy):
    """Host-array sgemv_hip / dgemv_hip (by y's dtype); y updated in place."""
    L = load()
    fn = L.dgemv_hip if y.dtype == np.float64 else L.sgemv_hip
    fn(_b(trans), M, N, alpha, _ptr(A), lda, _ptr(x), incx, beta, _ptr(y), incy)
    return _raise_if(L.ugemm_hip_last_error())


def axpy_dev(N, alpha, dx, incx, dy, incy, double=False, stream=None):
    s = stream.handle if hasattr(stream, "handle") else stream
    L = load()
    fn = L.ugemm_hip_daxpy_dev if double else L.ugemm_hip_saxpy_dev
    return _raise_if(fn(N, alpha, _ptr(dx), incx, _ptr(dy), incy, s))


def gemv_dev(trans, M, N, alpha, dA, lda, dx, incx, beta, dy, incy, double=False, stream=None):
    s = stream.handle if hasattr(stream, "handle") else stream
    L = load()
    fn = L.ugemm_hip_dgemv_dev if double else L.ugemm_hip_sgemv_dev
    return _raise_if(fn(_b(trans), M, N, alpha, _ptr(dA), lda, _ptr(dx), incx, beta, _ptr(dy), incy, s))


def timing_enable(on=True):
    """Bracket the core kernel of every following call (this thread) with HIP events."""
    load().ugemm_hip_timing_enable(1 if on else 0)


def mfma_ceiling(d_operands, fp16, rounds=20000):
    """Sustained 16x16x32 MFMA rate in TFLOP/s (f16 if fp16 else bf16) on the
    1024 16-byte operand fragments at device address d_operands."""
    tf = ctypes.c_double(0.0)
    _raise_if(load().ugemm_hip_mfma_ceiling(ctypes.c_void_p(d_operands), 1 if fp16 else 0, int(rounds),
                                           ctypes.byref(tf)))
    return tf.value


def timing_collect():
    """(number of timed launches, summed milliseconds) since the last collect."""
    ms = ctypes.c_double(0.0)
    n = load().ugemm_hip_timing_collect(ctypes.byref(ms))
    if n < 0:
        _raise_if(n)
    return n, ms.value


def variants(mode):
    L = load()
    return [L.ugemm_hip_variant_name(mode, v).decode() for v in range(L.ugemm_hip_variant_count(mode))]


def grouped_variants(mode=None):
    """Names of the grouped kernel's tile variants (indices of ugemm_hip_tune_grouped);
    with a mode, those of a plan of that mode."""
    L = load()
    if mode is None:
        return [L.ugemm_hip_grouped_variant_name(v).decode() for v in range(L.ugemm_hip_grouped_variant_count())]
    n = L.ugemm_hip_grouped_mode_variant_count(mode)
    if n < 0:
        raise UgemmError(n, f"grouped plans do not provide mode {mode}")
    return [L.ugemm_hip_grouped_mode_variant_name(mode, v).decode() for v in range(n)]


def force_variant(mode, variant):
    return _raise_if(load().ugemm_hip_force_variant(mode, variant))


def force_splits(mode, splits):
    """Split-K count of this thread's next calls (mode -1: DGEMM); 0 = automatic."""
    return _raise_if(load().ugemm_hip_force_splits(mode, splits))


class GroupedPlan:
    """A planned grouped SGEMM (one persistent launch, optionally a hipGraph).
    A list of ConvProblem makes a plan of convolutions instead."""

    def __init__(self, problems, mode=MODE_FP32, capture=True, stream=None):
        h = ctypes.c_void_p()
        s = stream.handle if hasattr(stream, "handle") else stream
        if problems and isinstance(problems[0], ConvProblem):
            self._arr = (ConvProblem * len(problems))(*problems)
            rc = load().ugemm_hip_conv_grouped_plan_create_mode(ctypes.byref(h), len(problems), self._arr, mode,
                                                               1 if capture else 0, s)
        else:
            self._arr = (Problem * len(problems))(*problems)
            rc = load().ugemm_hip_grouped_plan_create(ctypes.byref(h), len(problems), self._arr, mode,
                                                      1 if capture else 0, s)
        _raise_if(rc)
        self.handle = h.value

    def run(self, stream=None):
        s = stream.handle if hasattr(stream, "handle") else stream
        return _raise_if(load().ugemm_hip_grouped_run(self.handle, s))

    def tune(self):
        rc = load().ugemm_hip_tune_grouped(self.handle)
        if rc < 0:
            _raise_if(rc)
        return rc

    def destroy(self):
        if self.handle:
            load().ugemm_hip_grouped_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
