"""ctypes mirror of include/pcbgpu.h and a numpy-backed search problem that builds a pcbgpu_query.

The field layout follows the C header exactly; `Problem` keeps every array it hands out alive for as long as the
query is in use (the back end borrows them for the duration of a call, like GpuQuery in src/gpu.h:10-36).
"""
import ctypes as C

import numpy as np

INF = 0x3FFFFFFF
MAX_LAYERS = 8
CF_MAX = 0xFFFFF
W_MAX = 1023
FREE = 0
BLOCK = 0xFFFF
ERR_FAIL, ERR_CHAIN, ERR_BUSY, ERR_BUDGET = -1, -2, -3, -4
NOFIELD = 1

# move codes: 0 +x, 1 -x, 2 +y, 3 -y, 4 (+,+), 5 (+,-), 6 (-,+), 7 (-,-)  (src/router.cpp:1230)
DX = (1, -1, 0, 0, 1, 1, -1, -1)
DY = (0, 0, 1, -1, 1, -1, 1, -1)


class Entry(C.Structure):
    _fields_ = [("state", C.c_uint32), ("value", C.c_int32)]


class Query(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nl", C.c_int32),
        ("wi0", C.c_int32), ("wj0", C.c_int32), ("ww", C.c_int32), ("wh", C.c_int32),
        ("me", C.c_uint16),
        ("via_cost", C.c_int32),
        ("dirw", (C.c_int32 * 8) * MAX_LAYERS),
        ("lay", C.c_uint8 * MAX_LAYERS),
        ("cst", C.POINTER(C.c_int32) * MAX_LAYERS),
        ("st", C.POINTER(C.c_uint16) * MAX_LAYERS),
        ("vcst", C.POINTER(C.c_int32)),
        ("vst", C.POINTER(C.c_uint16)),
        ("tbit", C.POINTER(C.c_uint64)),
        ("own_cf", C.POINTER(Entry)), ("n_own_cf", C.c_size_t),
        ("own_vcf", C.POINTER(Entry)), ("n_own_vcf", C.c_size_t),
        ("prior", C.POINTER(C.c_uint8)),
        ("pi0", C.c_int32), ("pj0", C.c_int32), ("pw", C.c_int32), ("ph", C.c_int32), ("pnl", C.c_int32),
        ("pcol", C.POINTER(C.c_int32)), ("prow", C.POINTER(C.c_int32)),
        ("play", C.c_int32 * MAX_LAYERS),
        ("mlbias", C.c_int32 * 256),
        ("mthr", C.c_uint8), ("corridor", C.c_uint8),
        ("starts", C.POINTER(Entry)), ("n_starts", C.c_size_t),
    ]


class Out(C.Structure):
    _fields_ = [
        ("rx0", C.c_int32), ("ry0", C.c_int32), ("rw", C.c_int32), ("rh", C.c_int32),
        ("nl", C.c_int32), ("ww", C.c_int32), ("wh", C.c_int32),
        ("best", C.c_int64),
        ("d", C.POINTER(C.c_int32)),
        ("cap", C.c_size_t),
        ("sweeps", C.c_int32), ("tries", C.c_int32),
        ("ms_upload", C.c_double), ("ms_solve", C.c_double), ("ms_total", C.c_double),
        ("tile_visits", C.c_uint64),
    ]


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def entries(pairs):
    """(state, value) pairs -> a ctypes Entry array (or None)."""
    pairs = list(pairs)
    if not pairs:
        return None
    arr = (Entry * len(pairs))()
    for k, (s, v) in enumerate(pairs):
        arr[k].state = int(s)
        arr[k].value = int(v)
    return arr


def default_dirw(nl, orth=12, diag=17):
    w = np.zeros((MAX_LAYERS, 8), np.int32)
    w[:nl, :4] = orth
    w[:nl, 4:] = diag
    return w


class Problem:
    """One search: the full grid's maps plus the window, the net, the weights, starts and targets.

    cst: (nl, ny, nx) int32 cost maps (>= INF blocked, < 0 some net's own copper); st: (nl, ny, nx) uint16 static
    owners; vcst / vst: (ny, nx) the via maps. starts: [(window state, g0)]; targets: window states.
    """

    def __init__(self, cst, st, vcst, vst, *, starts, targets, window=None, me=0, via_cost=600, dirw=None,
                 lay=None, own_cf=(), own_vcf=(), prior=None):
        self.cst = np.ascontiguousarray(cst, np.int32)
        self.st = np.ascontiguousarray(st, np.uint16)
        self.vcst = np.ascontiguousarray(vcst, np.int32)
        self.vst = np.ascontiguousarray(vst, np.uint16)
        self.nl, self.ny, self.nx = self.cst.shape
        assert 1 <= self.nl <= MAX_LAYERS
        assert self.st.shape == self.cst.shape and self.vcst.shape == (self.ny, self.nx) == self.vst.shape
        self.window = tuple(window) if window is not None else (0, 0, self.nx, self.ny)
        wi0, wj0, ww, wh = self.window
        self.me = int(me)
        self.via_cost = int(via_cost)
        self.dirw = np.ascontiguousarray(default_dirw(self.nl) if dirw is None else dirw, np.int32).reshape(MAX_LAYERS, 8)
        self.lay = np.zeros(MAX_LAYERS, np.uint8)
        self.lay[:self.nl] = 1 if lay is None else np.asarray(lay, np.uint8)[:self.nl]
        self.starts = [(int(s), int(g)) for s, g in starts]
        self.targets = [int(t) for t in targets]
        ws = ww * wh * self.nl
        self.tbit = np.zeros(ws // 64 + 1, np.uint64)
        for t in self.targets:
            self.tbit[t >> 6] |= np.uint64(1) << np.uint64(t & 63)
        self.own_cf = [(int(s), int(v)) for s, v in own_cf]
        self.own_vcf = [(int(s), int(v)) for s, v in own_vcf]
        self.prior = prior  # dict(data (pnl, ph, pw) uint8, pi0, pj0, pcol (nx,), prow (ny,), play (8,), mlbias (256,), mthr, corridor)
        self._q = None

    def query(self):
        """The pcbgpu_query (built once; the Problem keeps its arrays alive)."""
        if self._q is not None:
            return self._q
        q = Query()
        q.nx, q.ny, q.nl = self.nx, self.ny, self.nl
        q.wi0, q.wj0, q.ww, q.wh = self.window
        q.me = self.me
        q.via_cost = self.via_cost
        for l in range(MAX_LAYERS):
            for d in range(8):
                q.dirw[l][d] = int(self.dirw[l, d])
            q.lay[l] = int(self.lay[l])
            q.play[l] = -1
        self._layers = [np.ascontiguousarray(self.cst[l]) for l in range(self.nl)]
        self._slayers = [np.ascontiguousarray(self.st[l]) for l in range(self.nl)]
        for l in range(self.nl):
            q.cst[l] = _ptr(self._layers[l], C.c_int32)
            q.st[l] = _ptr(self._slayers[l], C.c_uint16)
        q.vcst = _ptr(self.vcst, C.c_int32)
        q.vst = _ptr(self.vst, C.c_uint16)
        q.tbit = _ptr(self.tbit, C.c_uint64)
        self._own = entries(self.own_cf)
        self._ownv = entries(self.own_vcf)
        self._starts = entries(self.starts)
        if self._own is not None:
            q.own_cf, q.n_own_cf = self._own, len(self.own_cf)
        if self._ownv is not None:
            q.own_vcf, q.n_own_vcf = self._ownv, len(self.own_vcf)
        if self._starts is not None:
            q.starts, q.n_starts = self._starts, len(self.starts)
        if self.prior is not None:
            p = self.prior
            self._pdata = np.ascontiguousarray(p["data"], np.uint8)
            self._pcol = np.ascontiguousarray(p["pcol"], np.int32)
            self._prow = np.ascontiguousarray(p["prow"], np.int32)
            q.prior = _ptr(self._pdata, C.c_uint8)
            q.pnl, q.ph, q.pw = self._pdata.shape
            q.pi0, q.pj0 = int(p.get("pi0", 0)), int(p.get("pj0", 0))
            q.pcol = _ptr(self._pcol, C.c_int32)
            q.prow = _ptr(self._prow, C.c_int32)
            for l in range(MAX_LAYERS):
                q.play[l] = int(p["play"][l]) if l < len(p["play"]) else -1
            for v in range(256):
                q.mlbias[v] = int(p["mlbias"][v])
            q.mthr = int(p.get("mthr", 0))
            q.corridor = int(bool(p.get("corridor", False)))
        self._q = q
        return q

    def state(self, l, i, j):
        """Window state of window cell (i, j) on layer l."""
        wi0, wj0, ww, wh = self.window
        return (l * wh + j) * ww + i
