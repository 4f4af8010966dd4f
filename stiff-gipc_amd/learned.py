"""Learned coarse level: Python handles over the C ABI.

``Mpnn`` is the message-passing network over the MAS level-1 aggregate graph
(BF16 MFMA kernels on gfx950); ``MpnnData`` is a training set over one
level-1 graph.  The reference has no learned component; the model, its loss
and the training are described in ``include/sgipc.h``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import Context, DevArray, check, lib

__all__ = ["Mpnn", "MpnnData", "CHANNELS"]

CHANNELS = 64


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ptr)


def n_params(layers: int) -> int:
    return 3 * CHANNELS + layers * 2 * CHANNELS * CHANNELS + CHANNELS


class Mpnn:
    """Network parameters on the device; ``linear`` selects the identity
    activation instead of relu.  Uses the context it was made on (closing the
    context closes the network first)."""

    def __init__(self, ctx: Context, layers: int = 3, linear: bool = False, seed: int = 0):
        self.ctx, self.layers, self.linear = ctx, layers, bool(linear)
        h = ctypes.c_void_p()
        check(lib().sgipc_mpnn_create(ctx.h, layers, int(self.linear), ctypes.byref(h)), ctx.h)
        self.h = h
        ctx.adopt(self)
        if seed:
            check(lib().sgipc_mpnn_init_params(self.h, seed), ctx.h)

    @property
    def num_params(self) -> int:
        return lib().sgipc_mpnn_num_params(self.h)

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        check(lib().sgipc_mpnn_get_params(self.h, out.ctypes.data_as(ctypes.c_void_p), out.size), self.ctx.h)
        return out

    def set_params(self, v) -> None:
        v = np.ascontiguousarray(v, np.float32)
        check(lib().sgipc_mpnn_set_params(self.h, v.ctypes.data_as(ctypes.c_void_p), v.size), self.ctx.h)

    def forward(self, data: "MpnnData", first: int = 0, count: int = 1) -> np.ndarray:
        out = self.ctx.empty(count * 3 * data.nodes, np.float32)
        check(lib().sgipc_mpnn_forward(self.h, data.h, first, count, _p(out)), self.ctx.h)
        return out.download()

    def loss_grad(self, data: "MpnnData", first: int = 0, count: int = 1):
        loss = ctypes.c_double()
        g = np.zeros(self.num_params, np.float32)
        check(lib().sgipc_mpnn_loss_grad(self.h, data.h, first, count, ctypes.byref(loss),
                                         g.ctypes.data_as(ctypes.c_void_p)), self.ctx.h)
        return loss.value, g

    def train(self, data: "MpnnData", epochs: int, batch: int = 16, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        hist = np.zeros(epochs, np.float32)
        check(lib().sgipc_mpnn_train(self.h, data.h, epochs, batch, ctypes.c_float(lr), seed,
                                     hist.ctypes.data_as(ctypes.c_void_p)), self.ctx.h)
        return hist

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:  # a closed context closed this network first
                lib().sgipc_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MpnnData:
    """Training samples over one level-1 graph.  Either built from the upper
    BCOO of A_1 (``from_bcoo``) or filled by a system's collector.  The set
    uses the context it was made on (closing the context closes it first)."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h
        ctx.adopt(self)

    @classmethod
    def from_bcoo(cls, ctx: Context, n_nodes: int, vals, rows, cols):
        dv = ctx.to_device(np.ascontiguousarray(vals, np.float64))
        dr = ctx.to_device(np.ascontiguousarray(rows, np.int32), np.int32)
        dc = ctx.to_device(np.ascontiguousarray(cols, np.int32), np.int32)
        h = ctypes.c_void_p()
        check(lib().sgipc_mpnn_data_create(ctx.h, n_nodes, _p(dv), _p(dr), _p(dc), int(rows.size), ctypes.byref(h)),
              ctx.h)
        return cls(ctx, h)

    @property
    def nodes(self) -> int:
        return lib().sgipc_mpnn_data_nodes(self.h)

    @property
    def size(self) -> int:
        return lib().sgipc_mpnn_data_size(self.h)

    def add(self, X, q) -> None:
        X = np.ascontiguousarray(X, np.float32).reshape(-1)
        q = np.ascontiguousarray(q, np.float32).reshape(-1)
        ns = q.size // (3 * self.nodes)
        assert X.size == 3 * q.size and ns * 3 * self.nodes == q.size
        dX, dq = self.ctx.to_device(X, np.float32), self.ctx.to_device(q, np.float32)
        check(lib().sgipc_mpnn_data_add(self.h, _p(dX), _p(dq), ns), self.ctx.h)

    def samples(self, first: int = 0, count: int | None = None):
        count = self.size - first if count is None else count
        rows = 3 * self.nodes
        X = np.zeros(count * rows * 3, np.float32)
        q = np.zeros(count * rows, np.float32)
        check(lib().sgipc_mpnn_data_samples(self.h, first, count, X.ctypes.data_as(ctypes.c_void_p),
                                            q.ctypes.data_as(ctypes.c_void_p)), self.ctx.h)
        return X.reshape(count, rows, 3), q.reshape(count, rows)

    def graph(self):
        """(row_ptr, col, blk (nnz, 3, 3) row-major, scale) as built on the device."""
        n, nnz = self.nodes, lib().sgipc_mpnn_data_graph_nnz(self.h)
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        blk = np.zeros(9 * nnz, np.float32)
        sc = np.zeros(n, np.float32)
        vp = ctypes.c_void_p
        check(lib().sgipc_mpnn_data_graph(self.h, rp.ctypes.data_as(vp), col.ctypes.data_as(vp),
                                          blk.ctypes.data_as(vp), sc.ctypes.data_as(vp)), self.ctx.h)
        return rp, col, blk.reshape(nnz, 3, 3), sc

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:  # a closed context closed this set first
                lib().sgipc_mpnn_data_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
