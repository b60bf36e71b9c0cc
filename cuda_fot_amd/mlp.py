"""The learned cost model's network over libfot: an FP32 MLP
in -> hidden -> hidden -> out (ReLU) whose forward, backward and Adam steps
run on the GPU in libfot's hand-written kernels (matrix products on MFMA).
No reference counterpart."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import call, lib, ptr


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    def __init__(self, n_in: int = 24, hidden: int = 64, n_out: int = 6, seed: int = 0):
        h = C.c_void_p()
        call("fot_mlp_create", n_in, hidden, n_out, seed, C.byref(h))
        self.handle = h.value
        self.n_in, self.hidden, self.n_out = n_in, hidden, n_out
        n = C.c_uint64()
        call("fot_mlp_num_params", self.handle, C.byref(n))
        self.num_params = n.value

    # flat parameters: W1[h][in], b1, W2[h][h], b2, W3[out][h], b3
    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params, np.float32)
        call("fot_mlp_get_params", self.handle, ptr(p), p.size)
        return p

    def set_params(self, p) -> None:
        p = _f32(p)
        call("fot_mlp_set_params", self.handle, ptr(p), p.size)

    def unpack(self, p=None) -> dict:
        p = self.get_params() if p is None else p
        i, h, o = self.n_in, self.hidden, self.n_out
        sizes = [("W1", (h, i)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)), ("W3", (o, h)),
                 ("b3", (o,))]
        out, k = {}, 0
        for name, shape in sizes:
            n = int(np.prod(shape))
            out[name] = p[k:k + n].reshape(shape)
            k += n
        return out

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        y = np.empty((x.shape[0], self.n_out), np.float32)
        call("fot_mlp_forward", self.handle, ptr(x), x.shape[0], ptr(y))
        return y

    def loss_grad(self, x, t, mask=None):
        x, t = _f32(x), _f32(t)
        m = None if mask is None else _f32(mask)
        loss = C.c_float()
        g = np.empty(self.num_params, np.float32)
        call("fot_mlp_loss_grad", self.handle, ptr(x), ptr(t), ptr(m), x.shape[0],
             C.byref(loss), ptr(g))
        return loss.value, g

    def train(self, x, t, mask=None, epochs: int = 1000, lr: float = 3e-3,
              weight_decay: float = 0.0) -> np.ndarray:
        x, t = _f32(x), _f32(t)
        m = None if mask is None else _f32(mask)
        losses = np.empty(epochs, np.float32)
        call("fot_mlp_train", self.handle, ptr(x), ptr(t), ptr(m), x.shape[0], epochs, lr,
             weight_decay, ptr(losses))
        return losses

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.fot_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gemm(a: np.ndarray, b: np.ndarray, trans_a=False, trans_b=False) -> np.ndarray:
    """op(a) @ op(b) on the GPU through the MFMA GEMM (device round trip)."""
    from .device import DeviceArray
    a, b = _f32(a), _f32(b)
    M, K = (a.shape[1], a.shape[0]) if trans_a else a.shape
    N = b.shape[0] if trans_b else b.shape[1]
    da, db = DeviceArray.from_host(a), DeviceArray.from_host(b)
    dc = DeviceArray((M, N), np.float32)
    call("fot_gemm_f32", da.ptr, db.ptr, dc.ptr, M, N, K, int(trans_a), int(trans_b), None)
    from .device import synchronize
    synchronize()
    return dc.download()
