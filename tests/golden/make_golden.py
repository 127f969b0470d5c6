#!/usr/bin/env python3
"""Writes tests/golden/ref_gemm.npz: input/output vectors of the reference's
own CPU GEMM functions, produced by calling the reference library compiled
from its sources (oracle/_ref/libugemm_ref.so, built by oracle/Makefile from
ugemm.h where the reference is mounted).  Run here, where the reference
exists; the fixture is data (inputs and outputs), committed so that tests on
machines without the reference can pin the oracle and the GPU path to it.

Functions captured (reference file:line):
  sgemm_cpu   ugemm.h:287-411   all 8 major/trans combinations
  sgemm_c     gemm_cpu.h:284    all 8
  sgemm_avx   sgemm_avx256.h:392  R/N/N (its other layouts leave C untouched:
                                  captured for two of them to document that)
  sgemm_sse   sgemm_sse.h:365   all 8 (file-static; exported by oracle/ref_shim.c)
  dgemm_cpu   ugemm.h:162-286   all 8
  _dgemm_c    gemm_cpu.h:284    all 8
  dgemm_avx   dgemm_avx.h:844   row-major (4 combinations): its column-major
                                  store path (rs_c == 1, dgemm_avx.h:352-464)
                                  uses aligned 256-bit loads/stores on C and
                                  faults unless ldc*8 is a multiple of 32
Shapes: 1x1x1, 3x3x2 (the known answer of sgemm_test.c:186-200), 7x5x9 with
alpha in {0, 1, 1.5} x beta in {0, 0.5, 1}; 17x17x17 and 21x37x19 with
(alpha, beta) in {(0, .5), (1, 0), (1.5, .5), (1, 1)}; and the
reference's default check shape 128x361x1152 (check_sgemm.c:147-149) for
R/N/N with (alpha, beta) = (1, 0) and (1.5, 0.5).
Inputs: U[-1, 1) on 24 bits from numpy PCG64 (seed per case); stored for the
small shapes, regenerated from the seed (checked by SHA-256) for the large one.
"""
import hashlib
import json
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

FUNCS_S = ["sgemm_cpu", "sgemm_c", "sgemm_avx", "ref_sgemm_sse"]
FUNCS_D = ["dgemm_cpu", "_dgemm_c", "dgemm_avx"]
COMBOS = [(m, a, b) for m in "RC" for a in "NT" for b in "NT"]
TINY = [(1, 1, 1), (3, 3, 2), (7, 5, 9)]
SMALL = [(17, 17, 17), (21, 37, 19)]
AB = [(a, b) for a in (0.0, 1.0, 1.5) for b in (0.0, 0.5, 1.0)]
AB_SMALL = [(0.0, 0.5), (1.0, 0.0), (1.5, 0.5), (1.0, 1.0)]


def lds(major, ta, tb, M, N, K):
    row = major == "R"
    la = (M if ta == "T" else K) if row else (K if ta == "T" else M)
    lb = (K if tb == "T" else N) if row else (N if tb == "T" else K)
    lc = N if row else M
    return max(la, 1), max(lb, 1), max(lc, 1)


def main():
    L = orc.ref_lib()
    if L is None:
        sys.exit("oracle/_ref/libugemm_ref.so is not built (make -C oracle ref)")
    cases, arrays = [], {}

    inputs_of = {}

    def add(fn, dtype, major, ta, tb, M, N, K, alpha, beta, known=None):
        key = (np.dtype(dtype).name, major, ta, tb, M, N, K, alpha, beta, known is not None)
        seed = zlib.crc32(repr(key).encode())
        lda, ldb, ldc = lds(major, ta, tb, M, N, K)
        g = orc.rng(seed)
        sa = orc.stored_shape(major, ta, M, K, lda)
        sb = orc.stored_shape(major, tb, K, N, ldb)
        sc = (M, ldc) if major == "R" else (N, ldc)
        if known is not None:
            A, B, C0 = known
        else:
            A = orc.uniform(g, sa, dtype=dtype)
            B = orc.uniform(g, sb, dtype=dtype)
            C0 = orc.uniform(g, sc, dtype=dtype)
        # C sits at the start of a larger canary-filled buffer: the alpha == 0
        # branch of gemm_cpu.h:308-323 (and its copies) indexes C column-major
        # whatever `major` says, so for row-major calls it can write up to
        # C[(M-1) + (N-1)*ldc], past the M x ldc matrix.  Such writes are
        # caught here and recorded instead of corrupting the heap.
        n_c = C0.size
        span = max(n_c, (M - 1) + (N - 1) * ldc + 1, (N - 1) + (M - 1) * ldc + 1) + 64
        canary = dtype(-7.25e20) if dtype == np.float32 else dtype(-7.25e200)
        buf = np.full(span, canary, dtype=dtype)
        buf[:n_c] = C0.ravel()
        getattr(L, fn)(major.encode(), ta.encode(), tb.encode(), M, N, K, alpha, A.ctypes.data, lda, B.ctypes.data,
                       ldb, beta, buf.ctypes.data, ldc)
        oob = bool(np.any(buf[n_c:] != canary))
        C = buf[:n_c].reshape(C0.shape).copy()
        i = len(cases)
        big = A.size + B.size > 1 << 16
        h = hashlib.sha256(A.tobytes() + B.tobytes() + C0.tobytes()).hexdigest()
        if big:
            # Large inputs are not stored: they are regenerated from the seed
            # (orc.uniform on PCG64) and identified by their SHA-256.
            inp = None
        else:
            # Inputs are shared by every function of the same dtype and key.
            inp = inputs_of.get(h)
            if inp is None:
                inp = f"in{len(inputs_of)}"
                inputs_of[h] = inp
                arrays[inp + "_A"], arrays[inp + "_B"], arrays[inp + "_C0"] = A, B, C0
        arrays[f"C{i}"] = C
        cases.append(dict(fn=fn, dtype=np.dtype(dtype).name, major=major, ta=ta, tb=tb, M=M, N=N, K=K, alpha=alpha,
                          beta=beta, lda=lda, ldb=ldb, ldc=ldc, writes_outside_C=oob, seed=seed, inputs=inp,
                          inputs_sha256=h))

    for fn in FUNCS_S + FUNCS_D:
        dtype = np.float32 if fn in FUNCS_S else np.float64
        if fn == "sgemm_avx":
            combos = [("R", "N", "N"), ("R", "T", "N"), ("R", "N", "T")]
        elif fn == "dgemm_avx":
            combos = [c for c in COMBOS if c[0] == "R"]
        else:
            combos = COMBOS
        for (major, ta, tb) in combos:
            for (M, N, K) in TINY:
                for (alpha, beta) in AB:
                    add(fn, dtype, major, ta, tb, M, N, K, alpha, beta)
            for (M, N, K) in SMALL:
                for (alpha, beta) in AB_SMALL:
                    add(fn, dtype, major, ta, tb, M, N, K, alpha, beta)
    # Known answer of sgemm_test.c:186-200 (A = B = 1..6, C := A*B).
    kA = np.arange(1, 7, dtype=np.float32).reshape(3, 2)
    kB = np.arange(1, 7, dtype=np.float32).reshape(2, 3)
    for fn in FUNCS_S:
        add(fn, np.float32, "R", "N", "N", 3, 3, 2, 1.0, 0.0, known=(kA, kB, np.zeros((3, 3), np.float32)))
    # The reference's default check shape, R/N/N.
    for fn in FUNCS_S:
        for (alpha, beta) in ((1.0, 0.0), (1.5, 0.5)):
            add(fn, np.float32, "R", "N", "N", 128, 361, 1152, alpha, beta)

    out = os.path.join(HERE, "ref_gemm.npz")
    np.savez_compressed(out, **arrays)
    meta = dict(source="reference ugemm.h compiled by oracle/Makefile (flags of the reference Makefile, "
                       "-march=native replaced by -mavx2 -mfma)", cases=cases,
                sha256=hashlib.sha256(open(out, "rb").read()).hexdigest())
    with open(os.path.join(HERE, "ref_gemm.json"), "w") as f:
        json.dump(meta, f, indent=0)
    print(f"{len(cases)} cases -> {out} ({os.path.getsize(out) / 1e6:.2f} MB)")


if __name__ == "__main__":
    main()
