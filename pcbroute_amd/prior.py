"""The learned corridor prior: a conv3x3 CNN run and trained with the BF16 matrix-core kernels of libpcbgpu.so
(include/pcbprior.h). It stands where pcbroute calls ONNX Runtime for its RouteNet prior (src/mlinfer.cpp:834-856).

Tensors are numpy float32, pixel-major: x (n, h, w, in_ch), target (n, h, w, out_ch), mask (n, h, w); w must be a
multiple of 64 (pad the crop). There is no CPU path: every call raises if the library is missing.
"""
import ctypes as C

import numpy as np

from . import LIB_PATH, BackendError, load_library

MAX_CH = 32

_F = C.POINTER(C.c_float)


class Stats(C.Structure):
    _fields_ = [("conv_launches", C.c_uint64), ("wgrad_launches", C.c_uint64), ("conv_ms", C.c_double),
                ("wgrad_ms", C.c_double), ("conv_flop", C.c_double), ("wgrad_flop", C.c_double)]


def _bind(lib):
    lib.pcbprior_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]
    lib.pcbprior_create.restype = C.c_void_p
    lib.pcbprior_destroy.argtypes = [C.c_void_p]
    lib.pcbprior_num_params.argtypes = [C.c_void_p]
    lib.pcbprior_num_params.restype = C.c_size_t
    lib.pcbprior_get_params.argtypes = [C.c_void_p, _F]
    lib.pcbprior_set_params.argtypes = [C.c_void_p, _F]
    for f in (lib.pcbprior_predict, lib.pcbprior_logits):
        f.argtypes = [C.c_void_p, _F, C.c_int, C.c_int, C.c_int, _F]
    lib.pcbprior_predict_valid.argtypes = [C.c_void_p, _F, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int32), _F]
    lib.pcbprior_loss_grad.argtypes = [C.c_void_p, _F, _F, _F, C.c_int, C.c_int, C.c_int, _F, _F]
    lib.pcbprior_train_step.argtypes = [C.c_void_p, _F, _F, _F, C.c_int, C.c_int, C.c_int, C.c_float, _F]
    lib.pcbprior_mfma_check.argtypes = [_F, _F, _F]
    lib.pcbprior_dataset.argtypes = [C.c_void_p, _F, _F, _F, C.c_int, C.c_int, C.c_int]
    lib.pcbprior_train_resident.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int, C.c_float, _F]
    lib.pcbprior_forward_resident.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int, _F]
    lib.pcbprior_stats_enable.argtypes = [C.c_void_p, C.c_int]
    lib.pcbprior_stats_read.argtypes = [C.c_void_p, C.POINTER(Stats), C.c_int]
    lib.pcbprior_last_error.restype = C.c_char_p
    return lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
    return a


def _p(a):
    return a.ctypes.data_as(_F)


_lib = None


def library(path=LIB_PATH):
    global _lib
    if _lib is None:
        _lib = _bind(load_library(path))
    return _lib


def mfma_check(a, b):
    """One 16x16x32 BF16 matrix-core product a (16x32) @ b (32x16) (inputs rounded to BF16, FP32 result)."""
    lib = library()
    a, b = _f32(a, (16, 32)), _f32(b, (32, 16))
    d = np.zeros((16, 16), np.float32)
    if lib.pcbprior_mfma_check(_p(a), _p(b), _p(d)) != 0:
        raise BackendError(lib.pcbprior_last_error().decode())
    return d


class PriorNet:
    """`depth` conv3x3 layers, in_ch -> width -> ... -> out_ch, ReLU between layers, logits out."""

    def __init__(self, in_ch, width=32, depth=6, out_ch=4, seed=0):
        self.lib = library()
        self.in_ch, self.width, self.depth, self.out_ch = in_ch, width, depth, out_ch
        self.h = self.lib.pcbprior_create(in_ch, width, depth, out_ch, seed)
        if not self.h:
            raise BackendError(self.lib.pcbprior_last_error().decode())

    def close(self):
        if self.h:
            self.lib.pcbprior_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise BackendError(self.lib.pcbprior_last_error().decode() or "pcbprior call failed")

    @property
    def num_params(self):
        return int(self.lib.pcbprior_num_params(self.h))

    def layer_shapes(self):
        """[(w shape (co, 3, 3, ci), b shape (co,))] per layer, the order of get_params."""
        out = []
        for L in range(self.depth):
            ci = self.in_ch if L == 0 else self.width
            co = self.out_ch if L == self.depth - 1 else self.width
            out.append(((co, 3, 3, ci), (co,)))
        return out

    def get_params(self):
        v = np.zeros(self.num_params, np.float32)
        self._check(self.lib.pcbprior_get_params(self.h, _p(v)))
        return v

    def set_params(self, v):
        v = _f32(v, (self.num_params,))
        self._check(self.lib.pcbprior_set_params(self.h, _p(v)))

    def unpack(self, v):
        """The flat parameter vector as [(w (co, 3, 3, ci), b (co,))]."""
        out, o = [], 0
        for ws, bs in self.layer_shapes():
            nw = int(np.prod(ws))
            out.append((v[o:o + nw].reshape(ws), v[o + nw:o + nw + bs[0]]))
            o += nw + bs[0]
        return out

    def _x(self, x):
        x = _f32(x)
        if x.ndim != 4 or x.shape[3] != self.in_ch or x.shape[2] % 64:
            raise ValueError(f"x must be (n, h, w, {self.in_ch}) with w a multiple of 64, got {x.shape}")
        return x

    def logits(self, x):
        x = self._x(x)
        n, h, w, _ = x.shape
        out = np.zeros((n, h, w, self.out_ch), np.float32)
        self._check(self.lib.pcbprior_logits(self.h, _p(x), n, h, w, _p(out)))
        return out

    def predict(self, x):
        x = self._x(x)
        n, h, w, _ = x.shape
        out = np.zeros((n, h, w, self.out_ch), np.float32)
        self._check(self.lib.pcbprior_predict(self.h, _p(x), n, h, w, _p(out)))
        return out

    def predict_many(self, xs):
        """Images of different sizes (h, w, in_ch), w a multiple of 64, in one padded batch
        (pcbprior_predict_valid); each result equals predict on the image alone."""
        H, W = max(x.shape[0] for x in xs), max(x.shape[1] for x in xs)
        batch = np.zeros((len(xs), H, W, self.in_ch), np.float32)
        hw = np.zeros(2 * len(xs), np.int32)
        for k, x in enumerate(xs):
            batch[k, :x.shape[0], :x.shape[1]] = x
            hw[2 * k:2 * k + 2] = x.shape[:2]
        out = np.zeros((len(xs), H, W, self.out_ch), np.float32)
        self._check(self.lib.pcbprior_predict_valid(self.h, _p(batch), len(xs), H, W,
                                                    hw.ctypes.data_as(C.POINTER(C.c_int32)), _p(out)))
        return [out[k, :x.shape[0], :x.shape[1]] for k, x in enumerate(xs)]

    def _tm(self, x, target, mask):
        x = self._x(x)
        n, h, w, _ = x.shape
        return x, _f32(target, (n, h, w, self.out_ch)), _f32(mask, (n, h, w)), n, h, w

    def loss_grad(self, x, target, mask):
        x, t, m, n, h, w = self._tm(x, target, mask)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.num_params, np.float32)
        self._check(self.lib.pcbprior_loss_grad(self.h, _p(x), _p(t), _p(m), n, h, w, _p(loss), _p(g)))
        return float(loss[0]), g

    def dataset(self, x, target, mask):
        """Makes (x, target, mask) the resident training set (kept in HBM as BF16 inputs)."""
        x, t, m, n, h, w = self._tm(x, target, mask)
        self._check(self.lib.pcbprior_dataset(self.h, _p(x), _p(t), _p(m), n, h, w))
        self._ds = (n, h, w)

    def train_resident(self, idx, lr=1e-3, loss=True):
        """One Adam step on samples idx of the resident set; the loss (or None)."""
        idx = np.ascontiguousarray(idx, np.int32)
        out = np.zeros(1, np.float32)
        self._check(self.lib.pcbprior_train_resident(self.h, idx.ctypes.data_as(C.POINTER(C.c_int32)), len(idx), lr,
                                                     _p(out) if loss else None))
        return float(out[0]) if loss else None

    def forward_resident(self, idx, logits=True):
        idx = np.ascontiguousarray(idx, np.int32)
        n, h, w = self._ds
        out = np.zeros((len(idx), h, w, self.out_ch), np.float32) if logits else None
        self._check(self.lib.pcbprior_forward_resident(self.h, idx.ctypes.data_as(C.POINTER(C.c_int32)), len(idx),
                                                       _p(out) if logits else None))
        return out

    def stats(self, enable=None, reset=False):
        if enable is not None:
            self.lib.pcbprior_stats_enable(self.h, int(bool(enable)))
        st = Stats()
        self.lib.pcbprior_stats_read(self.h, C.byref(st), int(reset))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def train_step(self, x, target, mask, lr=1e-3):
        x, t, m, n, h, w = self._tm(x, target, mask)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.pcbprior_train_step(self.h, _p(x), _p(t), _p(m), n, h, w, lr, _p(loss)))
        return float(loss[0])
