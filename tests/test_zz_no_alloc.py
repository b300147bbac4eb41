"""sparc_sample_noalloc (staging in the library's fixed device array, no
device allocation) against sparc_sample (which allocates its staging) on the
same state, through the C ABI: bit-equal on two calls in a row, the second
with fewer points, and a point count above the maximum rejected."""
import ctypes as C

import numpy as np
import pytest

from sparc_amd import abi
from sparc_amd import driver as D

pytestmark = pytest.mark.gpu


def _points(shape, n, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    return np.stack([rng.integers(0, nx, n), rng.integers(0, ny, n),
                     rng.integers(1, nz + 1, n)], axis=1).astype(np.int32)


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32], ids=["fp64", "fp32"])
def test_sample_noalloc_matches_sample(prec):
    W = D.workload("tiny")
    S = D.setup_for(W)
    eng, cfg = D.make_engine(W, S, precision=prec, nsteps=6)
    try:
        mask, series = D.forward_source(W, 14)
        eng.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        rng = np.random.default_rng(3)
        for f in range(6):
            eng.put(abi.ARR_A, f, rng.standard_normal(eng.shape))
        eng.step(0, 3)
        for n, seed in ((5000, 1), (37, 2)):
            ijk = _points(eng.shape, n, seed)
            for f in (abi.V_Z, abi.XI_X):
                want = eng.sample(f, ijk)
                got = eng.sample(f, ijk, noalloc=True)
                assert np.array_equal(got, want), (n, f)
                a = eng.get(abi.ARR_A, f)
                if cfg.compensated:  # sampling reads the high FP32 word of the hi + 16-bit lo state
                    a = a - eng.get(abi.ARR_ALO, f)
                assert np.array_equal(got, a[ijk[:, 0], ijk[:, 1], ijk[:, 2] - 1])
        too_many = _points(eng.shape, 65537, 4)
        with pytest.raises(abi.SparcError):
            eng.sample(abi.V_Z, too_many, noalloc=True)
        # the context stays usable after the rejected call
        ijk = _points(eng.shape, 64, 5)
        assert np.array_equal(eng.sample(abi.V_Z, ijk, noalloc=True), eng.sample(abi.V_Z, ijk))
    finally:
        eng.close()
