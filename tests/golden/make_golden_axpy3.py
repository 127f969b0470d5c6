#!/usr/bin/env python3
"""Writes tests/golden/ref_axpy3.npz: input/output vectors of the reference's
three-operand axpy helpers, z = a*x + y, produced by calling the reference
library compiled from its sources (oracle/_ref/libugemm_ref.so, built by
oracle/Makefile from ugemm.h with -mfma, the FMA branch that the reference's
-march=native build takes).  Run here, where the reference exists; the
fixture is data only.

Functions captured (reference file:line):
  saxpy_avx  ugemm.h:56-73   N in {8, 64, 1000, 4096, 13}
  daxpy_avx  ugemm.h:41-55   N in {4, 64, 1000, 4096, 13}
The reference works in whole vectors of 8 (float) or 4 (double) elements, so
for N = 13 it reads and writes up to the next multiple: x and y are padded
to that length and the whole written z is recorded (key *_zfull).  Arrays
are 32-byte aligned (_mm256_load_ps/pd and _mm256_stream_ps/pd need it).
Inputs: U[-1, 1) on 24 bits from numpy PCG64 (seed per case); a in
{1.5, -0.7}.
"""
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402


def aligned(n, dt):
    """A zeroed 1-D array of n elements whose data starts on a 32-byte boundary."""
    item = np.dtype(dt).itemsize
    raw = np.zeros(n + 32 // item, dt)
    off = (-raw.ctypes.data % 32) // item
    return raw[off:off + n]


def main():
    R = orc.ref_lib()
    if R is None:
        sys.exit("oracle/_ref/libugemm_ref.so is not built (the reference is needed to make this fixture)")
    arrays, index = {}, []
    for dt, name, width, sizes in ((np.float32, "saxpy_avx", 8, (8, 64, 1000, 4096, 13)),
                                   (np.float64, "daxpy_avx", 4, (4, 64, 1000, 4096, 13))):
        for N in sizes:
            for a in (1.5, -0.7):
                key = f"{name}_{N}_{a}"
                g = orc.rng(zlib.crc32(key.encode()))
                n_full = (N + width - 1) // width * width
                x, y, z = aligned(n_full, dt), aligned(n_full, dt), aligned(n_full, dt)
                x[:] = orc.uniform(g, (n_full,), dtype=dt)
                y[:] = orc.uniform(g, (n_full,), dtype=dt)
                getattr(R, name)(N, a, orc._p(x), orc._p(y), orc._p(z))
                arrays[key + "_x"], arrays[key + "_y"] = x.copy(), y.copy()
                arrays[key + "_zfull"] = z.copy()
                index.append((key, name, N, a))
    arrays["index"] = np.array([repr(t) for t in index])
    out = os.path.join(HERE, "ref_axpy3.npz")
    np.savez_compressed(out, **arrays)
    print(f"{len(index)} cases -> {out} ({os.path.getsize(out)} bytes)")


if __name__ == "__main__":
    main()
