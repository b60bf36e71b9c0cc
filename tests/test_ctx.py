"""The context layer of the C-ABI (fot_ctx_*, fot_buf_*, fot_plan_*,
fot_loop_launch, fot_graph_*, fot_batch_launch): argument errors on the CPU;
on the GPU, the one-shot launch and a captured step replayed by hand give the
bits of the Loop path on the same inputs."""
import copy
import ctypes as C

import numpy as np
import pytest

from cuda_fot_amd import capi
from cuda_fot_amd.device import make_desc
from helpers import airfoil_problem, assert_bitwise

RES_CALC = capi.functor(capi.KERNEL_RES_CALC)


def test_context_argument_errors():
    lib = capi.lib
    assert lib.fot_ctx_create(None, 0) == capi.ERR_INVALID_ARGUMENT
    assert lib.fot_ctx_destroy(None) == capi.OK
    assert lib.fot_ctx_synchronize(None) == capi.ERR_INVALID_ARGUMENT
    assert lib.fot_ctx_stream(None, None) == capi.ERR_INVALID_ARGUMENT
    p = C.c_void_p()
    assert lib.fot_buf_alloc(None, 8, C.byref(p)) == capi.ERR_INVALID_ARGUMENT
    assert lib.fot_buf_free(None, None) == capi.ERR_INVALID_ARGUMENT
    assert lib.fot_plan_upload(None, None, None, None, None) == capi.ERR_INVALID_ARGUMENT
    assert lib.fot_loop_launch(None, None, None, None, 1) == capi.ERR_INVALID_ARGUMENT
    assert lib.fot_graph_capture(None, None, None) == capi.ERR_INVALID_ARGUMENT
    assert lib.fot_graph_launch(None, None) == capi.ERR_INVALID_ARGUMENT
    assert lib.fot_graph_free(None) == capi.OK
    assert lib.fot_batch_launch(None, None, 1) == capi.ERR_INVALID_ARGUMENT
    if capi.device_count() < 1:
        ctx = C.c_void_p()
        assert lib.fot_ctx_create(C.byref(ctx), 0) == capi.ERR_NO_DEVICE


class _Buf:
    """A context-owned device copy of a host array, shaped for make_desc."""

    def __init__(self, ctx, host):
        host = np.ascontiguousarray(host)
        self.shape = host.shape
        p = C.c_void_p()
        capi.call("fot_buf_alloc", ctx, host.nbytes, C.byref(p))
        self.ptr = p.value
        capi.call("fot_buf_upload", ctx, self.ptr, capi.ptr(host), host.nbytes)

    def download(self, ctx, like):
        out = np.empty_like(like)
        capi.call("fot_buf_download", ctx, capi.ptr(out), self.ptr, out.nbytes)
        return out


def _desc(ctx, p):
    dev = [_Buf(ctx, d) for d in p.point_data]
    out = _Buf(ctx, p.point_data[0])
    cdev = [_Buf(ctx, c) for c in p.cell_data]
    d, keep = make_desc(p.maps, p.num_points, dev, out, cdev, soa=p.soa)
    return d, keep, dev, out


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, capi.LOOP_OWNER_COMPUTE | capi.LOOP_TWO_PER_THREAD |
                                   capi.LOOP_OWNER_GATHER])
def test_context_launch_and_graph_match_loop(gpu, flags):
    iters = 3
    base = airfoil_problem(nx=40, ny=24, res0_seed=31)
    ref = copy.deepcopy(base)
    ref.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, iters, block_size=128, flags=flags,
                 partition=False)
    ctx = C.c_void_p()
    capi.call("fot_ctx_create", C.byref(ctx), 0)
    try:
        cfg = capi.LoopConfig(capi.STRATEGY_HIERARCHICAL, 128, None, 0, flags)
        # one-shot: plan, run, free (the reference's loopGPUHierarchical call)
        d, keep, dev, out = _desc(ctx, base)
        capi.call("fot_loop_launch", ctx, C.byref(RES_CALC), C.byref(d), C.byref(cfg), iters)
        assert_bitwise(dev[0].download(ctx, base.point_data[0]), ref.point_data[0],
                       "fot_loop_launch vs Loop")
        # a kept plan, one step captured in a graph and replayed by hand with
        # fot_loop_run's copies around it
        d, keep, dev, out = _desc(ctx, base)
        plan = C.c_void_p()
        graph = C.c_void_p()
        capi.call("fot_plan_upload", ctx, C.byref(RES_CALC), C.byref(d), C.byref(cfg),
                  C.byref(plan))
        capi.call("fot_graph_capture", ctx, plan, C.byref(graph))
        s = C.c_void_p()
        capi.call("fot_ctx_stream", ctx, C.byref(s))
        nbytes = base.point_data[0].nbytes
        capi.call("fot_copy", out.ptr, dev[0].ptr, nbytes, s)
        for _ in range(iters):
            capi.call("fot_graph_launch", ctx, graph)
        capi.call("fot_copy", dev[0].ptr, out.ptr, nbytes, s)
        capi.call("fot_ctx_synchronize", ctx)
        assert_bitwise(dev[0].download(ctx, base.point_data[0]), ref.point_data[0],
                       "captured step replayed vs Loop")
        capi.call("fot_graph_free", graph)
        capi.call("fot_plan_free", plan)
        # a buffer of another context is refused
        other = C.c_void_p()
        capi.call("fot_ctx_create", C.byref(other), 0)
        assert capi.lib.fot_buf_free(other, C.c_void_p(dev[0].ptr)) == capi.ERR_INVALID_ARGUMENT
        capi.call("fot_ctx_destroy", other)
        capi.call("fot_buf_free", ctx, C.c_void_p(dev[0].ptr))
    finally:
        capi.call("fot_ctx_destroy", ctx)


@pytest.mark.gpu
def test_context_graph_of_an_empty_plan_is_a_no_op(gpu):
    base = airfoil_problem(nx=8, ny=6, res0_seed=32)
    base.maps = [m[:0].copy() for m in base.maps]
    ctx = C.c_void_p()
    capi.call("fot_ctx_create", C.byref(ctx), 0)
    try:
        d, keep, dev, out = _desc(ctx, base)
        cfg = capi.LoopConfig(capi.STRATEGY_HIERARCHICAL, 128, None, 0, 0)
        plan, graph = C.c_void_p(), C.c_void_p()
        capi.call("fot_plan_upload", ctx, C.byref(RES_CALC), C.byref(d), C.byref(cfg),
                  C.byref(plan))
        capi.call("fot_graph_capture", ctx, plan, C.byref(graph))
        capi.call("fot_graph_launch", ctx, graph)
        capi.call("fot_loop_launch", ctx, C.byref(RES_CALC), C.byref(d), C.byref(cfg), 2)
        capi.call("fot_ctx_synchronize", ctx)
        assert_bitwise(dev[0].download(ctx, base.point_data[0]), base.point_data[0],
                       "empty plan")
        capi.call("fot_graph_free", graph)
        capi.call("fot_plan_free", plan)
    finally:
        capi.call("fot_ctx_destroy", ctx)
