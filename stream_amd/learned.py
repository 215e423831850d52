"""Learned L L^T preconditioner: a message-passing network trained on the
GPU by the library's own forward and backward kernels.

No counterpart in the reference, whose only preconditioner is PrecJacobi
(include/numerics/preconditioner.hpp:60-72, python/numerics/jacobi.py);
`PrecLearned` follows that class's create/destroy pattern and is consumed
by `PfemStep2D.run(solver="learned")` and `PfemStep3D.run(solver="learned")`.
Node features are (x, y, z, boundary flag, A_ii / 8), z = 0 for 2D
systems, so one architecture serves both.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import StreamError, check, lib
from ._ct import arr, bind, i32, opt, u32

_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
PrecLearned_create = bind(lib, "PrecLearned_create", _p, [])
PrecLearned_destroy = bind(lib, "PrecLearned_destroy", None, [_p])
PrecLearned_param_count = bind(lib, "PrecLearned_param_count", u32, [])
PrecLearned_set_params = bind(lib, "PrecLearned_set_params", i32, [_p, arr(np.float32, 1), u32])
PrecLearned_get_params = bind(lib, "PrecLearned_get_params", i32, [_p, arr(np.float32, 1, writeable=True), u32])
PrecLearned_set_adam = bind(lib, "PrecLearned_set_adam", i32, [_p, ctypes.c_float, ctypes.c_float, ctypes.c_float])
PrecLearned_set_system = bind(lib, "PrecLearned_set_system", i32,
                              [_p, u32, arr(np.uint32, 1), arr(np.uint32, 1), arr(np.float64, 1),
                               arr(np.float64, 2), arr(np.uint8, 1)])
PrecLearned_set_system_dim = bind(lib, "PrecLearned_set_system_dim", i32,
                                  [_p, u32, u32, arr(np.uint32, 1), arr(np.uint32, 1), arr(np.float64, 1),
                                   arr(np.float64, 2), arr(np.uint8, 1)])
PrecLearned_bind_step = bind(lib, "PrecLearned_bind_step", i32, [_p, _p])
PrecLearned_bind_step3d = bind(lib, "PrecLearned_bind_step3d", i32, [_p, _p])
PrecLearned_forward = bind(lib, "PrecLearned_forward", i32, [_p, _dp, opt(np.float32)])
PrecLearned_gradient = bind(lib, "PrecLearned_gradient", i32, [_p, _dp, opt(np.float32)])
PrecLearned_train_step = bind(lib, "PrecLearned_train_step", i32, [_p, ctypes.c_float, _dp])
PrecLearned_time_gemm = bind(lib, "PrecLearned_time_gemm", i32, [_p, u32, _dp, _dp])

HIDDEN = 64
LAYERS = 3
NODE_FEATS = 5
# Parameter count of files written before the z feature existed (4 node
# features); load_params widens them with a zero z row, which leaves every
# 2D output unchanged.
_COUNT_4FEAT = 63168
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def param_layout():
    """(name, shape) of the parameter tensors in flat order ([in][out] matrices)."""
    H = HIDDEN
    out = [("enc_w", (NODE_FEATS, H)), ("enc_b", (H,))]
    for l in range(LAYERS):
        out += [(f"pq_w{l}", (H, 2 * H)), (f"e_w{l}", (3, H)), (f"m_b{l}", (H,)), (f"u_w{l}", (2 * H, H)),
                (f"u_b{l}", (H,))]
    out += [("dec_w", (H, 3 * H)), ("de_w", (3, H)), ("de_b", (H,)), ("o_w", (H,)), ("dd_b", (H,)), ("d_w", (H,))]
    return out


def init_params(seed: int = 7) -> np.ndarray:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights from SplitMix64,
    zero biases and zero output weights (the initial factor is Jacobi)."""
    from .clouds import uniform

    n = int(sum(np.prod(s) for _, s in param_layout()))
    u = uniform(seed, n)
    flat = np.zeros(n)
    o = 0
    for name, shape in param_layout():
        k = int(np.prod(shape))
        is_w = name.endswith("_w") or name[:-1].endswith("_w")
        if is_w and name not in ("o_w", "d_w"):
            flat[o:o + k] = (2.0 * u[o:o + k] - 1.0) / np.sqrt(shape[0])
        o += k
    return flat.astype(np.float32)


class PrecLearned:
    def __init__(self, params: np.ndarray | None = None):
        self.h = PrecLearned_create()
        if not self.h:
            raise StreamError(lib.Stream_last_error().decode())
        self.count = int(PrecLearned_param_count())
        self.nnz = 0
        self.set_params(init_params() if params is None else params)

    def __del__(self):
        if getattr(self, "h", None):
            PrecLearned_destroy(self.h)
            self.h = None

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.shape != (self.count,):
            raise ValueError(f"expected {self.count} parameters, got {w.shape}")
        check(PrecLearned_set_params(self.h, w, self.count))

    def params(self) -> np.ndarray:
        w = np.empty(self.count, np.float32)
        check(PrecLearned_get_params(self.h, w, self.count))
        return w

    def set_adam(self, beta1=0.9, beta2=0.999, eps=1e-8):
        check(PrecLearned_set_adam(self.h, beta1, beta2, eps))

    def set_system(self, row, col, val, xy, bnd):
        """Host system; xy holds 2 or 3 coordinates per row."""
        row = np.ascontiguousarray(row, np.uint32)
        n = row.shape[0] - 1
        xy = np.ascontiguousarray(xy, np.float64)
        check(PrecLearned_set_system_dim(self.h, xy.shape[1], n, row, np.ascontiguousarray(col, np.uint32),
                                         np.ascontiguousarray(val, np.float64), xy,
                                         np.ascontiguousarray(bnd, np.uint8)))
        self.nnz = int(row[-1])

    def bind_step(self, step):
        """Point the network at a PfemStep2D / PfemStep3D's assembled system."""
        from .pfem import PfemStep3D

        if isinstance(step, PfemStep3D):
            check(PrecLearned_bind_step3d(self.h, step.h))
        else:
            check(PrecLearned_bind_step(self.h, step.h))
        self.nnz = step.nnz()

    def forward(self, with_ghat=False):
        loss = ctypes.c_double(0.0)
        g = np.empty(self.nnz, np.float32) if with_ghat else None
        check(PrecLearned_forward(self.h, ctypes.byref(loss), g))
        return (loss.value, g) if with_ghat else loss.value

    def gradient(self):
        loss = ctypes.c_double(0.0)
        g = np.empty(self.count, np.float32)
        check(PrecLearned_gradient(self.h, ctypes.byref(loss), g))
        return loss.value, g

    def train_step(self, lr: float, read_loss: bool = True):
        loss = ctypes.c_double(0.0)
        check(PrecLearned_train_step(self.h, lr, ctypes.byref(loss) if read_loss else None))
        return loss.value if read_loss else None

    def time_gemm(self, reps: int = 100):
        """(average ms, FLOPs) of the layer-0 MFMA GEMM on the bound system."""
        ms = ctypes.c_double(0.0)
        fl = ctypes.c_double(0.0)
        check(PrecLearned_time_gemm(self.h, reps, ctypes.byref(ms), ctypes.byref(fl)))
        return ms.value, fl.value

    def save(self, path: str, meta: dict | None = None):
        """FP32 parameters (little-endian .bin) plus a JSON description."""
        w = self.params()
        w.astype("<f4").tofile(path + ".bin")
        info = dict(count=self.count, hidden=HIDDEN, layers=LAYERS,
                    layout=[[n, list(s)] for n, s in param_layout()])
        info.update(meta or {})
        with open(path + ".json", "w") as f:
            json.dump(info, f, indent=1)

    @staticmethod
    def load_params(path: str) -> np.ndarray:
        with open(path + ".json") as f:
            info = json.load(f)
        w = np.fromfile(path + ".bin", dtype="<f4")
        if w.shape[0] != info["count"]:
            raise ValueError("weight file does not match its description")
        if w.shape[0] == _COUNT_4FEAT:
            w = np.concatenate([w[:2 * HIDDEN], np.zeros(HIDDEN, np.float32), w[2 * HIDDEN:]])
        return w.astype(np.float32)
