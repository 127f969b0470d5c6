"""The sustained matrix-core rate probe (ugemm_hip_mfma_ceiling): a positive
rate below the published dense peak for both element types, argument checks."""
import numpy as np
import pytest

import bench
import ugemm_amd as ug
from ugemm_amd import hip

pytestmark = pytest.mark.gpu


def test_mfma_ceiling_rates():
    rates = bench.mfma_sustained(rounds=2000)
    for name in ("bf16", "f16"):
        # 16 MFMAs x 16384 flop per round on every wave: a rate between a
        # tenth of the published dense peak (2516.6 TFLOP/s) and the peak.
        assert 250.0 < rates[name] <= 2516.6 * 1.02, rates


def test_mfma_ceiling_arguments():
    d = hip.DeviceArray.from_host(np.zeros(8192, np.uint16))
    try:
        with pytest.raises(ug.UgemmError) as e:
            ug.mfma_ceiling(d.ptr, True, 0)
        assert e.value.status == ug.ERR_ARG
        with pytest.raises(ug.UgemmError) as e:
            ug.mfma_ceiling(0, False, 10)
        assert e.value.status == ug.ERR_ARG
    finally:
        d.free()
