"""Device arithmetic against the reference's own scalar helpers and IEEE fp32 (bit-exact)."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def test_device_binning_and_distance_bits(dev):
    z = np.load(helpers.GOLDEN / "arith.npz", allow_pickle=False)
    v = z["v"]
    rng = np.random.Generator(np.random.PCG64(3))
    w = rng.permutation(v).astype(np.float32)
    # denormal products and tiny distances must survive (no flush to zero)
    w[:2000] = rng.uniform(1e-30, 1e-19, 2000).astype(np.float32)
    o = dev.selftest_arith(v, w)
    for k in ("xbin", "ybin", "eta_bin", "phi_bin"):
        assert np.array_equal(o[k], z[k]), k
    assert np.array_equal(o["reduce"].view(np.int32), z["reduce"].view(np.int32))
    with np.errstate(all="ignore"):
        dist = np.sqrt((v * v).astype(np.float32) + (w * w).astype(np.float32)).astype(np.float32)
        quot = (v / w).astype(np.float32)
    assert np.array_equal(o["dist"].view(np.int32), dist.view(np.int32))
    ok = ~np.isnan(quot)
    assert np.array_equal(o["quot"][ok].view(np.int32), quot[ok].view(np.int32))
