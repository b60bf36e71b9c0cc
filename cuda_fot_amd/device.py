"""Device buffers, streams, timers and the loop engine over libfot."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import call, lib, ptr


class DeviceArray:
    """A HIP device allocation holding a numpy-shaped array."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        call("fot_malloc", C.byref(p), self.nbytes)
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch {a.nbytes} != {self.nbytes}")
        call("fot_memcpy_htod", self.ptr, ptr(a), self.nbytes)

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        call("fot_memcpy_dtoh", ptr(out), self.ptr, self.nbytes)
        return out

    def free(self) -> None:
        if self.ptr:
            lib.fot_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    call("fot_device_synchronize")


class Stream:
    def __init__(self):
        s = C.c_void_p()
        call("fot_stream_create", C.byref(s))
        self.handle = s.value

    def synchronize(self) -> None:
        call("fot_stream_synchronize", self.handle)

    def __del__(self):
        try:
            if self.handle:
                lib.fot_stream_destroy(self.handle)
        except Exception:
            pass


class Timer:
    """HIP-event timer on a stream (replaces CudaTimer, timer.hpp:70-102)."""

    def __init__(self):
        t = C.c_void_p()
        call("fot_timer_create", C.byref(t))
        self.handle = t.value

    def start(self, stream=None) -> None:
        call("fot_timer_start", self.handle, stream)

    def stop(self, stream=None) -> None:
        call("fot_timer_stop", self.handle, stream)

    def elapsed_ms(self) -> float:
        ms = C.c_float()
        call("fot_timer_elapsed_ms", self.handle, C.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                lib.fot_timer_destroy(self.handle)
        except Exception:
            pass


def make_desc(maps, num_points, point_data, out, cell_data=(), soa=False):
    """A fot_problem_desc over host maps and device point / cell data;
    returns (desc, keepalive)."""
    maps = [capi.u32(m) for m in maps]
    d = capi.ProblemDesc()
    d.num_cells = maps[0].shape[0]
    d.num_mappings = len(maps)
    for k, m in enumerate(maps):
        d.cell_to_node[k] = ptr(m)
        d.mesh_dim[k] = m.shape[1]
        d.num_points[k] = int(num_points[k])
        pd = point_data[k]
        d.point_data[k] = pd.ptr if pd is not None else None
        d.point_dim[k] = pd.shape[1] if pd is not None and len(pd.shape) > 1 else 1
        if soa and pd is not None and len(pd.shape) > 1:
            d.point_dim[k] = pd.shape[0]
    d.num_cell_data = len(cell_data)
    for k, cd in enumerate(cell_data):
        d.cell_data[k] = cd.ptr
        d.cell_dim[k] = cd.shape[0] if len(cd.shape) > 1 else 1
    d.point_data_out = out.ptr
    d.soa = 1 if soa else 0
    # the device arrays too: the desc holds their raw addresses, and a
    # DeviceArray frees its memory when the last reference goes
    return d, (maps, list(point_data), out, tuple(cell_data))


class Batch:
    """Many problems of one functor run by one persistent launch per step
    (fot_batch).  Each problem: dict(maps, num_points, point_data, out,
    cell_data, soa, block_size, point_partition=None) with device arrays."""

    def __init__(self, func: capi.Functor, problems, use_graph=False):
        n = len(problems)
        descs = (capi.ProblemDesc * n)()
        cfgs = (capi.LoopConfig * n)()
        self._keep = []
        for i, pr in enumerate(problems):
            d, keep = make_desc(pr["maps"], pr["num_points"], pr["point_data"], pr["out"],
                                pr.get("cell_data", ()), pr.get("soa", False))
            descs[i] = d
            self._keep.append(keep)
            cfgs[i].strategy = capi.STRATEGY_HIERARCHICAL
            cfgs[i].block_size = int(pr.get("block_size", 128))
            part = pr.get("point_partition")
            if part is not None:
                part = capi.u32(part)
                self._keep.append(part)
                cfgs[i].partition = ptr(part)
                cfgs[i].flags = capi.LOOP_POINT_PARTITION
        h = C.c_void_p()
        call("fot_batch_create", C.byref(func), n, descs, cfgs, 1 if use_graph else 0,
             C.byref(h))
        self.handle = h.value

    def step(self, stream=None) -> None:
        call("fot_batch_step", self.handle, stream)

    def run(self, iters: int, stream=None) -> None:
        call("fot_batch_run", self.handle, iters, stream)

    def stats(self) -> dict:
        s = capi.BatchStats()
        call("fot_batch_stats_get", self.handle, C.byref(s))
        return {f: getattr(s, f) for f, _ in capi.BatchStats._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.fot_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Loop:
    """A planned loop (fot_loop): strategy, plan on the device, optional graph.

    `maps` are host (num_cells, mesh_dim) u32 arrays, one per mapping;
    `point_data` device arrays (point_data[0] is the iterate and is updated by
    run()); `out` the device accumulation buffer.
    """

    def __init__(self, func: capi.Functor, maps, num_points, point_data, out,
                 cell_data=(), soa=False, strategy=capi.STRATEGY_HIERARCHICAL,
                 block_size=256, partition=None, use_graph=False, flags=0):
        self._maps = [capi.u32(m) for m in maps]
        self._partition = None if partition is None else capi.u32(partition)
        # the loop launches kernels on these buffers' raw addresses: hold them
        # for as long as the loop exists, so a caller dropping its own
        # references cannot free memory the loop still reads and writes
        self._buffers = (list(point_data), out, tuple(cell_data))
        d = capi.ProblemDesc()
        d.num_cells = self._maps[0].shape[0]
        d.num_mappings = len(self._maps)
        for k, m in enumerate(self._maps):
            d.cell_to_node[k] = ptr(m)
            d.mesh_dim[k] = m.shape[1]
            d.num_points[k] = int(num_points[k])
            d.point_data[k] = point_data[k].ptr if point_data[k] is not None else None
            d.point_dim[k] = (point_data[k].shape[1] if point_data[k] is not None and
                              len(point_data[k].shape) > 1 else 1)
            if soa and point_data[k] is not None and len(point_data[k].shape) > 1:
                d.point_dim[k] = point_data[k].shape[0]
        d.num_cell_data = len(cell_data)
        for k, cd in enumerate(cell_data):
            d.cell_data[k] = cd.ptr
            d.cell_dim[k] = cd.shape[0] if len(cd.shape) > 1 else 1
        d.point_data_out = out.ptr
        d.soa = 1 if soa else 0
        self.desc = d
        cfg = capi.LoopConfig()
        cfg.strategy = strategy
        cfg.block_size = block_size
        cfg.partition = ptr(self._partition)
        cfg.use_graph = 1 if use_graph else 0
        cfg.flags = flags
        self.cfg = cfg
        h = C.c_void_p()
        call("fot_loop_create", C.byref(func), C.byref(d), C.byref(cfg), C.byref(h))
        self.handle = h.value

    def step(self, stream=None) -> None:
        call("fot_loop_step", self.handle, stream)

    def run(self, iters: int, stream=None) -> None:
        call("fot_loop_run", self.handle, iters, stream)

    def stats(self) -> dict:
        s = capi.LoopStats()
        call("fot_loop_stats_get", self.handle, C.byref(s))
        return {f: getattr(s, f) for f, _ in capi.LoopStats._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.fot_loop_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
