"""Small host-to-device uploads go through a 32 MiB pinned ring that
synchronises the device and starts over when it fills.  Forty box-copy calls
with ~1 MiB task tables each (14000 single-element boxes, 64 + 8 bytes per
box) push more than the ring's size through it; every call's scattered copy
(alternately negated) must land exactly."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_upload_ring_wraps(gpu):
    print("case upload ring: 40 calls x 14000 boxes", flush=True)
    rng = np.random.default_rng(11)
    n = 14000
    x = rng.standard_normal(n)
    src = gpu.DeviceArray.from_numpy(x)
    dst = gpu.DeviceArray(n, np.float64)
    lib = gpu.lib()
    for call in range(40):
        perm = rng.permutation(n)
        sign = -1.0 if call % 2 else 1.0
        tasks = np.zeros((n, 10), dtype=np.int64)
        tasks[:, 0] = np.arange(n)
        tasks[:, 3] = perm
        tasks[:, 6:9] = 1
        tasks[:, 9] = -1 if sign < 0 else 0
        gpu.check(lib.dmrgx_copy_boxes(gpu.R64, src.ptr, dst.ptr, tasks.ctypes.data_as(C.POINTER(C.c_int64)), n),
                  "dmrgx_copy_boxes")
        expect = np.empty(n)
        expect[perm] = sign * x
        assert np.array_equal(dst.numpy(), expect), call
    src.free()
    dst.free()
