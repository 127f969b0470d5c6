#!/usr/bin/env python3
"""Writes tests/golden/ref_blas12.npz: input/output vectors of the reference's
BLAS-1/2 helpers, produced by calling the reference library compiled from
its sources (oracle/_ref/libugemm_ref.so, built by oracle/Makefile from
ugemm.h where the reference is mounted).  Run here, where the reference
exists; the fixture is data only.

Functions captured (reference file:line):
  saxpy_cpu / daxpy_cpu  ugemm.h:76-101   N in {1, 7, 64, 1000}, (incx, incy) in {(1,1), (2,1), (1,3), (3,2)}
  sgemv_cpu / dgemv_cpu  ugemm.h:103-160  trans N/T, (M, N) in {(1,1), (5,7), (33,17), (64,129)},
                                          lda = rows + {0, 3}, incx == incy in {1, 2} with
                                          (alpha, beta) in {(1, 0), (1.5, 0.5), (0, 2)};
                                          plus incx != incy cases, where the reference reads x
                                          with incy (ugemm.h:116,147): flagged x_by_incy.
Inputs: U[-1, 1) on 24 bits from numpy PCG64 (seed per case).
"""
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402


def main():
    R = orc.ref_lib()
    if R is None:
        sys.exit("oracle/_ref/libugemm_ref.so is not built (the reference is needed to make this fixture)")
    arrays, index = {}, []
    for dt, name in ((np.float32, "saxpy_cpu"), (np.float64, "daxpy_cpu")):
        for N in (1, 7, 64, 1000):
            for incx, incy in ((1, 1), (2, 1), (1, 3), (3, 2)):
                key = f"{name}_{N}_{incx}_{incy}"
                g = orc.rng(zlib.crc32(key.encode()))
                x = orc.uniform(g, ((N - 1) * incx + 1,), dtype=dt)
                y0 = orc.uniform(g, ((N - 1) * incy + 1,), dtype=dt)
                y = y0.copy()
                alpha = 1.5
                getattr(R, name)(N, alpha, orc._p(x), incx, orc._p(y), incy)
                arrays[key + "_x"], arrays[key + "_y0"], arrays[key + "_y"] = x, y0, y
                index.append((key, name, N, incx, incy, alpha, 0.0, "", 0, 0, 0))
    for dt, name in ((np.float32, "sgemv_cpu"), (np.float64, "dgemv_cpu")):
        for trans in "NT":
            for M, N in ((1, 1), (5, 7), (33, 17), (64, 129)):
                for pad in (0, 3):
                    cases = [(1, 1, a, b) for a, b in ((1.0, 0.0), (1.5, 0.5), (0.0, 2.0))]
                    cases += [(2, 2, 1.5, 0.5), (1, 2, 1.5, 0.5), (3, 1, 1.0, 0.0)]
                    for incx, incy, alpha, beta in cases:
                        key = f"{name}_{trans}_{M}_{N}_{pad}_{incx}_{incy}_{alpha}_{beta}"
                        g = orc.rng(zlib.crc32(key.encode()))
                        rows, cols = (M, N) if trans == "N" else (N, M)
                        lda = rows + pad
                        A = orc.uniform(g, (cols * lda,), dtype=dt)
                        # the reference reads x with incy: size x for both readings
                        nx = (N - 1) * max(incx, incy) + 1
                        x = orc.uniform(g, (nx,), dtype=dt)
                        y0 = orc.uniform(g, ((M - 1) * incy + 1,), dtype=dt)
                        y = y0.copy()
                        getattr(R, name)(trans.encode(), M, N, alpha, orc._p(A), lda, orc._p(x), incx, beta,
                                         orc._p(y), incy)
                        arrays[key + "_A"], arrays[key + "_x"] = A, x
                        arrays[key + "_y0"], arrays[key + "_y"] = y0, y
                        index.append((key, name, N, incx, incy, alpha, beta, trans, M, lda, int(incx != incy)))
    arrays["index"] = np.array([repr(t) for t in index])
    out = os.path.join(HERE, "ref_blas12.npz")
    np.savez_compressed(out, **arrays)
    print(f"{len(index)} cases -> {out} ({os.path.getsize(out)} bytes)")


if __name__ == "__main__":
    main()
