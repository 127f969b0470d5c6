"""Shared helpers for the parity tests: build a seeded GEMM case, run it
through libugemm_hip.so (host- or device-pointer entry), and judge it against
the CPU oracle."""
import numpy as np

from oracle import oracle as orc

# Extra per-product error terms (relative to |a||b|) of the reduced-precision
# operand forms, on top of the FP32 accumulation bound orc.tolerance(K):
#   bf16x3: dropped lo*lo and split residuals, |r| <= 2^-16 |x|  -> 3 * 2^-16 (+ slack)
#   bf16x6: dropped terms of order 2^-24                         -> 8 * 2^-24
#   bf16x1: both operands rounded to 8 significant bits         -> 2^-7 + 2^-16 (+ slack)
#   fp16x3: fp16 split residuals and lo*lo, each <= 2^-22 |a||b|  -> 3 * 2^-22 (+ slack)
MODE_EXTRA = {0: 0.0, 1: 2.0 ** -14, 2: 2.0 ** -6, 4: 2.0 ** -21, 5: 2.0 ** -19}


class Case:
    def __init__(self, major, ta, tb, M, N, K, alpha=1.0, beta=0.0, seed=1, lda=None, ldb=None, ldc=None,
                 dist="uniform", dtype=np.float32, pad=0):
        self.major, self.ta, self.tb = major, ta, tb
        self.M, self.N, self.K = M, N, K
        self.alpha, self.beta = alpha, beta
        self.dtype = np.dtype(dtype)
        row = major == "R"
        need_a = (M if ta != "N" else K) if row else (K if ta != "N" else M)
        need_b = (K if tb != "N" else N) if row else (N if tb != "N" else K)
        need_c = N if row else M
        self.lda = lda or max(1, need_a + pad)
        self.ldb = ldb or max(1, need_b + pad)
        self.ldc = ldc or max(1, need_c + pad)
        g = orc.rng(seed)
        gen = {"uniform": orc.uniform, "normal": orc.normal, "wide": orc.wide_range}[dist]
        sa = orc.stored_shape(major, ta, M, K, self.lda)
        sb = orc.stored_shape(major, tb, K, N, self.ldb)
        sc = (M, self.ldc) if row else (N, self.ldc)
        self.A = gen(g, sa, dtype=self.dtype)
        self.B = gen(g, sb, dtype=self.dtype)
        self.C0 = gen(g, sc, dtype=self.dtype)

    def exact(self, rows=None):
        return orc.reference_f64(self.major, self.ta, self.tb, self.M, self.N, self.K, self.alpha, self.A,
                                 self.lda, self.B, self.ldb, self.beta, self.C0, self.ldc, rows=rows)

    def logical(self, C):
        """C[m, n] view of a result buffer."""
        return orc.c_matrix(self.major, C, self.M, self.N)

    def judge(self, C, mode=0, eps=orc.EPS32):
        exact, scale = self.exact()
        if self.major != "R":
            pass  # reference_f64 already indexes C[m, n]
        tol = orc.tolerance(self.K, eps=eps, extra=MODE_EXTRA.get(mode, 0.0))
        return orc.check_bound(self.logical(C), exact, scale, tol)

    def untouched_padding(self, C):
        """True when the gaps between rows (ld > row length) kept C0's values."""
        row_len = self.N if self.major == "R" else self.M
        return np.array_equal(C[:, row_len:], self.C0[:, row_len:])
