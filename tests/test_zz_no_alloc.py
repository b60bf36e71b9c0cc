"""fot_mlp_forward_noalloc: the cost model's inference with its hidden
activations in the library's fixed per-device array instead of allocated
buffers.  Checked through the C-ABI against fot_mlp_forward (the same
kernels in the same order: identical bits) on two calls in a row, the first
at the array's largest row count and the second smaller, and on a size the
array cannot hold, which must be refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from cuda_fot_amd import capi
from cuda_fot_amd.device import DeviceArray, Stream, synchronize
from cuda_fot_amd.mlp import MLP

SCRATCH_FLOATS = 1 << 20  # the array's size in floats (include/fot_capi.h)


@pytest.mark.gpu
def test_mlp_forward_noalloc_matches_forward(gpu):
    n_in, hidden, n_out = 24, 64, 6
    net = MLP(n_in, hidden, n_out, seed=7)
    rng = np.random.default_rng(11)
    max_rows = SCRATCH_FLOATS // (2 * hidden)
    stream = Stream()
    for call_no, (rows, s) in enumerate([(max_rows, None), (700, stream.handle)]):
        print(f"noalloc call {call_no}: rows {rows}", flush=True)
        x = rng.normal(size=(rows, n_in)).astype(np.float32)
        want = net.forward(x)
        dx = DeviceArray.from_host(x)
        dy = DeviceArray.from_host(np.full((rows, n_out), np.nan, np.float32))
        capi.call("fot_mlp_forward_noalloc", net.handle, dx.ptr, rows, dy.ptr, s)
        if s is None:
            synchronize()
        else:
            stream.synchronize()
        got = dy.download()
        np.testing.assert_array_equal(got, want, err_msg=f"call {call_no}, rows {rows}")

    # one row more than the array holds: refused with a status, no launch
    dx = DeviceArray((1,), np.float32)
    st = capi.lib.fot_mlp_forward_noalloc(C.c_void_p(net.handle), C.c_void_p(dx.ptr),
                                         C.c_uint32(max_rows + 1), C.c_void_p(dx.ptr), None)
    assert st == capi.ERR_INVALID_ARGUMENT
    assert "scratch" in capi.last_error()
    # zero rows is a no-op
    capi.call("fot_mlp_forward_noalloc", net.handle, dx.ptr, 0, dx.ptr, None)
    net.close()
