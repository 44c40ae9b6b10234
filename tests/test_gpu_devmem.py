"""Library-owned device memory (msbg_dev_*): round trips and a convolution
on DeviceArray tensors against the same convolution through the library."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M = pytest.importorskip("msbg_amd")


def test_device_array_round_trip():
    rng = np.random.default_rng(5)
    for shape, dt in [((7,), np.float32), ((3, 5, 2), np.uint16), ((1,), np.float64)]:
        a = (rng.random(shape) * 1000).astype(dt)
        with M.DeviceArray.from_numpy(a) as d:
            assert d.data_ptr() != 0
            assert np.array_equal(d.numpy(), a)
            b = (rng.random(shape) * 1000).astype(dt)
            d.upload(b)
            assert np.array_equal(d.numpy(), b)
    z = M.DeviceArray((0,), np.float32)
    assert z.data_ptr() == 0
    z.free()
