"""The GPU path against the reference's own outputs (tests/golden/ref_gemm.npz):
for every golden case of the reference's naive loops (sgemm_cpu, dgemm_cpu)
and of its AVX/packed paths where they follow BLAS semantics, the GPU result
(plain FP32 SGEMM or FP64 DGEMM, host-pointer entry) must lie within
2(K+1) eps * scale of the reference's result, the reference itself being
within the same bound of the exact product."""
import numpy as np
import pytest

import ugemm_amd as ug
from oracle import oracle as orc
from tests.test_oracle import golden_inputs, load_golden, reference_divergence

pytestmark = pytest.mark.gpu


def test_gpu_matches_reference_golden_vectors():
    cases, data = load_golden()
    n = 0
    for i, c in enumerate(cases):
        if reference_divergence(c) is not None:
            continue
        A, B, C0 = golden_inputs(c, data)
        C = C0.copy()
        args = (c["major"], c["ta"], c["tb"], c["M"], c["N"], c["K"], c["alpha"], A, c["lda"], B, c["ldb"],
                c["beta"], C, c["ldc"])
        if c["dtype"] == "float32":
            ug.sgemm(*args)
            eps = orc.EPS32
        else:
            ug.dgemm(*args)
            eps = orc.EPS64
        _, scale = orc.reference_f64(c["major"], c["ta"], c["tb"], c["M"], c["N"], c["K"], c["alpha"], A, c["lda"],
                                     B, c["ldb"], c["beta"], C0, c["ldc"])
        ref = orc.c_matrix(c["major"], data[f"C{i}"], c["M"], c["N"]).astype(np.float64)
        # Both results lie within tol*scale of the exact product: they differ by at most twice that.
        ok, worst = orc.check_bound(orc.c_matrix(c["major"], C, c["M"], c["N"]), ref, scale,
                                    2 * orc.tolerance(c["K"], eps=eps))
        assert ok, (c, worst)
        n += 1
    assert n > 1000
