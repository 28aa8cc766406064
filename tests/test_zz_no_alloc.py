"""lbm_mlp_predict_noalloc, the allocation-free variant of lbm_mlp_predict
(include/lbm_mlp.h), through the C ABI: the same results bit for bit on two
calls in a row, the second with fewer samples (the device array keeps the
first call's contents beyond them), for a BF16 and an FP32 ensemble sharing
the GPU's array one call at a time; more samples than the array holds are
rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _predict(lib, h, x, noalloc):
    n = x.shape[0]
    mean = np.full((n, 6), np.nan, np.float32)
    spread = np.full(n, np.nan, np.float32)
    fn = lib.lbm_mlp_predict_noalloc if noalloc else lib.lbm_mlp_predict
    rc = fn(h, _fp(x), n, _fp(mean), _fp(spread))
    assert rc == 0, lib.lbm_hip_last_error().decode()
    return mean, spread


@pytest.mark.parametrize("mode,nets", [("bf16", 4), ("fp32", 3)])
def test_predict_noalloc_equals_predict(lbm, mode, nets):
    rng = np.random.default_rng(11)
    with lbm.MLP(nets=nets, mode=mode, seed=5) as m:
        lib = m.lib
        for n in (3000, 700):  # the second call is smaller
            x = np.ascontiguousarray(rng.uniform(-1, 1, size=(n, 5)).astype(np.float32))
            ma, sa = _predict(lib, m.h, x, False)
            mb, sb = _predict(lib, m.h, x, True)
            assert np.isfinite(ma).all() and np.isfinite(sa).all()
            assert np.array_equal(ma, mb)
            assert np.array_equal(sa, sb)
        # the full-size call and one past it
        x = np.ascontiguousarray(rng.uniform(-1, 1, size=(4097, 5)).astype(np.float32))
        ma, sa = _predict(lib, m.h, x[:4096], False)
        mb, sb = _predict(lib, m.h, x[:4096], True)
        assert np.array_equal(ma, mb) and np.array_equal(sa, sb)
        mean = np.empty((4097, 6), np.float32)
        rc = lib.lbm_mlp_predict_noalloc(m.h, _fp(x), 4097, _fp(mean), None)
        assert rc == -1 and "at most 4096" in lib.lbm_hip_last_error().decode()
