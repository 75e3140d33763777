"""pcbroute_amd - the MI355X grid-search back end of pcbroute (HIP, gfx950), behind the C ABI of include/pcbgpu.h.

`Backend` binds libpcbgpu.so (built in-tree by `python -m pcbroute_amd.build`). There is no fallback: importing
works without the library, but every Backend call raises if it is missing.
"""
import ctypes as C
import os

import numpy as np

from .abi import (BLOCK, CF_MAX, DX, DY, ERR_BUDGET, ERR_BUSY, ERR_CHAIN, ERR_FAIL, FREE, INF, MAX_LAYERS, NOFIELD, W_MAX,
                  Entry, Out, Problem, Query, default_dirw)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpcbgpu.so")

EXPORTS = (
    "pcbgpu_available", "pcbgpu_configure", "pcbgpu_solve", "pcbgpu_out_release", "pcbgpu_out_at",
    "pcbgpu_backtrace", "pcbgpu_grid_create", "pcbgpu_grid_destroy", "pcbgpu_grid_upload", "pcbgpu_solve_grid",
    "pcbgpu_solve_rect_grid", "pcbgpu_round_begin", "pcbgpu_round_submit", "pcbgpu_round_submit_rect", "pcbgpu_round_flags", "pcbgpu_round_launch",
    "pcbgpu_round_result", "pcbgpu_round_end", "pcbgpu_set_tile", "pcbgpu_get_tile", "pcbgpu_set_delta", "pcbgpu_stats_enable",
    "pcbgpu_stats_read", "pcbgpu_last_error", "pcbgpu_autotune", "pcbgpu_tune_get", "pcbgpu_tune_set",
    "pcbgpu_tune_save", "pcbgpu_tune_load", "pcbgpu_grid_upload_rects", "pcbgpu_solve_grid_noalloc",
    "pcbgpu_cpu_search",
)


class Tune(C.Structure):
    _fields_ = [("tile_rows", C.c_int32), ("local_sweeps", C.c_int32), ("delta", C.c_int32), ("tried", C.c_int32),
                ("ms", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("solves", C.c_uint64), ("launches", C.c_uint64), ("passes", C.c_uint64), ("tiles", C.c_uint64),
                ("states", C.c_uint64), ("rect_state_passes", C.c_uint64), ("kernel_ms", C.c_double),
                ("solve_ms", C.c_double), ("state_sweeps", C.c_uint64)]


class BackendError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise BackendError(f"{path} is missing: build it with `python -m pcbroute_amd.build`")
    lib = C.CDLL(path)
    lib.pcbgpu_available.restype = C.c_int
    lib.pcbgpu_configure.argtypes = [C.c_int, C.c_size_t]
    lib.pcbgpu_solve.argtypes = [C.POINTER(Query), C.POINTER(Out), C.c_int]
    lib.pcbgpu_solve.restype = C.c_int
    lib.pcbgpu_out_release.argtypes = [C.POINTER(Out)]
    lib.pcbgpu_out_at.argtypes = [C.POINTER(Out), C.c_int64]
    lib.pcbgpu_out_at.restype = C.c_int32
    lib.pcbgpu_backtrace.argtypes = [C.POINTER(Query), C.POINTER(Out), C.c_int32, C.c_double, C.POINTER(C.c_uint32),
                                     C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.c_size_t]
    lib.pcbgpu_backtrace.restype = C.c_int64
    lib.pcbgpu_grid_create.argtypes = [C.c_int32, C.c_int32, C.c_int32]
    lib.pcbgpu_grid_create.restype = C.c_void_p
    lib.pcbgpu_grid_destroy.argtypes = [C.c_void_p]
    lib.pcbgpu_grid_upload_rects.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                             C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_uint16))]
    lib.pcbgpu_grid_upload_rects.restype = C.c_int
    lib.pcbgpu_grid_upload.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_uint16)]
    lib.pcbgpu_grid_upload.restype = C.c_int
    lib.pcbgpu_grid_upload_rect.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                            C.POINTER(C.c_int32), C.POINTER(C.c_uint16)]
    lib.pcbgpu_grid_upload_rect.restype = C.c_int
    lib.pcbgpu_solve_grid.argtypes = [C.c_void_p, C.POINTER(Query), C.POINTER(Out), C.c_int]
    lib.pcbgpu_solve_grid.restype = C.c_int
    lib.pcbgpu_solve_rect_grid.argtypes = [C.c_void_p, C.POINTER(Query), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                           C.POINTER(Out), C.c_int]
    lib.pcbgpu_solve_rect_grid.restype = C.c_int
    lib.pcbgpu_solve_grid_noalloc.argtypes = [C.c_void_p, C.POINTER(Query), C.POINTER(C.c_int32), C.c_size_t,
                                              C.POINTER(Out)]
    lib.pcbgpu_solve_grid_noalloc.restype = C.c_int
    lib.pcbgpu_cpu_search.argtypes = [C.POINTER(Query), C.c_int, C.POINTER(C.c_int32), C.c_size_t, C.c_int32, C.c_size_t,
                                      C.c_int32, C.c_double, C.POINTER(C.c_uint32), C.c_size_t,
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.c_size_t]
    lib.pcbgpu_cpu_search.restype = C.c_int64
    lib.pcbgpu_set_tile.argtypes = [C.c_int32, C.c_int32]
    lib.pcbgpu_set_tile.restype = C.c_int
    lib.pcbgpu_get_tile.argtypes = [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.pcbgpu_set_delta.argtypes = [C.c_int32]
    lib.pcbgpu_set_delta.restype = C.c_int
    lib.pcbgpu_round_begin.restype = C.c_void_p
    lib.pcbgpu_round_submit.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Query)]
    lib.pcbgpu_round_submit.restype = C.c_int
    lib.pcbgpu_round_submit_rect.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Query), C.c_int32, C.c_int32,
                                             C.c_int32, C.c_int32]
    lib.pcbgpu_round_submit_rect.restype = C.c_int
    lib.pcbgpu_round_flags.argtypes = [C.c_void_p, C.c_int]
    lib.pcbgpu_round_launch.argtypes = [C.c_void_p]
    lib.pcbgpu_round_launch.restype = C.c_int
    lib.pcbgpu_round_result.argtypes = [C.c_void_p, C.c_int, C.POINTER(Out)]
    lib.pcbgpu_round_result.restype = C.c_int
    lib.pcbgpu_round_end.argtypes = [C.c_void_p]
    lib.pcbgpu_stats_enable.argtypes = [C.c_int]
    lib.pcbgpu_stats_read.argtypes = [C.POINTER(Stats), C.c_int]
    lib.pcbgpu_last_error.restype = C.c_char_p
    lib.pcbgpu_autotune.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.POINTER(Query)), C.c_int, C.c_int,
                                    C.POINTER(Tune)]
    lib.pcbgpu_autotune.restype = C.c_int
    lib.pcbgpu_tune_get.argtypes = [C.c_int32, C.POINTER(Tune)]
    lib.pcbgpu_tune_get.restype = C.c_int
    lib.pcbgpu_tune_set.argtypes = [C.c_int32, C.POINTER(Tune)]
    lib.pcbgpu_tune_set.restype = C.c_int
    lib.pcbgpu_tune_save.argtypes = [C.c_char_p]
    lib.pcbgpu_tune_save.restype = C.c_int
    lib.pcbgpu_tune_load.argtypes = [C.c_char_p]
    lib.pcbgpu_tune_load.restype = C.c_int
    return lib


def cpu_search(prob, heuristic=0, boxes=None, max_pops=0, via_gapc=0, via_gap2=0.0, old_vias=(), cf_min=1, lib=None):
    """The CPU search (pcbgpu_cpu_search, host code: no device needed). Returns dict(rc, best, opt, pops, chain):
    rc > 0 the chain length with chain [(window state, code)] target first, 0 no target reachable, ERR_BUDGET the
    budget ran out, ERR_CHAIN the via rules rejected the chain. boxes: rows (l, i0, j0, i1, j1) or None."""
    lib = lib or load_library()
    WN = prob.window[2] * prob.window[3]
    cap = WN * prob.nl + 1
    states = np.zeros(cap, np.uint32)
    codes = np.zeros(cap, np.uint8)
    ov = np.asarray(list(old_vias), np.uint32)
    bx = np.ascontiguousarray(np.asarray(boxes if boxes is not None else np.zeros((0, 5)), np.int32).reshape(-1, 5))
    best, opt, pops = C.c_int64(), C.c_int32(), C.c_uint64()
    n = lib.pcbgpu_cpu_search(C.byref(prob.query()), int(heuristic),
                              bx.ctypes.data_as(C.POINTER(C.c_int32)) if len(bx) else None, len(bx), int(cf_min),
                              int(max_pops),
                              int(via_gapc), float(via_gap2), ov.ctypes.data_as(C.POINTER(C.c_uint32)), len(ov),
                              C.byref(best), C.byref(opt), C.byref(pops), states.ctypes.data_as(C.POINTER(C.c_uint32)),
                              codes.ctypes.data_as(C.POINTER(C.c_uint8)), cap)
    chain = [(int(states[k]), int(codes[k])) for k in range(n)] if n > 0 else []
    return dict(rc=int(n), best=int(best.value), opt=int(opt.value), pops=int(pops.value), chain=chain)


class Result:
    """A solve's output copied out of the pooled pinned buffer: rect (rx0, ry0, rw, rh), best window state,
    field d (nl, rh, rw) int32, sweeps (passes of the last rectangle), tries, timings."""

    def __init__(self, out, rc, copy_field=True):
        self.rc = rc
        self.rect = (out.rx0, out.ry0, out.rw, out.rh)
        self.nl, self.ww, self.wh = out.nl, out.ww, out.wh
        self.best = int(out.best)
        self.sweeps = int(out.sweeps)
        self.tries = int(out.tries)
        self.ms_upload, self.ms_solve, self.ms_total = out.ms_upload, out.ms_solve, out.ms_total
        self.tile_visits = int(out.tile_visits)
        self.d = None
        if copy_field and out.d:
            n = out.nl * out.rw * out.rh
            self.d = np.ctypeslib.as_array(out.d, shape=(n,)).copy().reshape(out.nl, out.rh, out.rw)

    def at(self, ws):
        """Field value of a window state (INF outside the solved rectangle)."""
        WN = self.ww * self.wh
        l, c = divmod(ws, WN)
        j, i = divmod(c, self.ww)
        rx0, ry0, rw, rh = self.rect
        i -= rx0
        j -= ry0
        if not (0 <= i < rw and 0 <= j < rh):
            return INF
        return int(self.d[l, j, i])


class Grid:
    """Device-resident copy of a board's maps (pcbgpu_grid)."""

    def __init__(self, backend, cst, st, vcst, vst):
        self.b = backend
        cst = np.ascontiguousarray(cst, np.int32)
        st = np.ascontiguousarray(st, np.uint16)
        self.nl, self.ny, self.nx = cst.shape
        self.h = backend.lib.pcbgpu_grid_create(self.nx, self.ny, self.nl)
        if not self.h:
            raise BackendError(backend.error())
        for l in range(self.nl + 1):
            c = np.ascontiguousarray(cst[l] if l < self.nl else vcst, np.int32)
            s = np.ascontiguousarray(st[l] if l < self.nl else vst, np.uint16)
            if backend.lib.pcbgpu_grid_upload(self.h, l, 0, self.ny, c.ctypes.data_as(C.POINTER(C.c_int32)),
                                              s.ctypes.data_as(C.POINTER(C.c_uint16))) != 0:
                raise BackendError(backend.error())

    def upload_rects(self, rects, cst, st, vcst, vst):
        """Rectangles (layer, i0, j0, w, h) of new maps (layer nl: the via grid) in one call
        (pcbgpu_grid_upload_rects)."""
        cst = np.ascontiguousarray(cst, np.int32)
        st = np.ascontiguousarray(st, np.uint16)
        vcst = np.ascontiguousarray(vcst, np.int32)
        vst = np.ascontiguousarray(vst, np.uint16)
        n = len(rects)
        flat = np.ascontiguousarray(np.asarray(rects, np.int32).reshape(n * 5))
        cs = (C.POINTER(C.c_int32) * max(1, n))()
        ss = (C.POINTER(C.c_uint16) * max(1, n))()
        for k, r in enumerate(rects):
            c, s = (cst[r[0]], st[r[0]]) if r[0] < self.nl else (vcst, vst)
            cs[k] = c.ctypes.data_as(C.POINTER(C.c_int32))
            ss[k] = s.ctypes.data_as(C.POINTER(C.c_uint16))
        if self.b.lib.pcbgpu_grid_upload_rects(self.h, n, flat.ctypes.data_as(C.POINTER(C.c_int32)), cs, ss) != 0:
            raise BackendError(self.b.error())

    def close(self):
        if self.h:
            self.b.lib.pcbgpu_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Backend:
    def __init__(self, path=LIB_PATH):
        self.lib = load_library(path)

    def error(self):
        e = self.lib.pcbgpu_last_error()
        return e.decode() if e else ""

    def available(self):
        return bool(self.lib.pcbgpu_available())

    def configure(self, slots, max_states=0):
        self.lib.pcbgpu_configure(int(slots), int(max_states))

    def set_tile(self, rows=0, local_sweeps=0):
        if self.lib.pcbgpu_set_tile(int(rows), int(local_sweeps)) != 0:
            raise BackendError("bad tile setting")

    def set_delta(self, delta):
        if self.lib.pcbgpu_set_delta(int(delta)) != 0:
            raise BackendError("bad delta")

    def stats(self, enable=None, reset=False):
        """Kernel accounting: enable / disable it, and read (and optionally reset) the counters."""
        if enable is not None:
            self.lib.pcbgpu_stats_enable(int(bool(enable)))
        st = Stats()
        self.lib.pcbgpu_stats_read(C.byref(st), int(reset))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def get_tile(self, nl):
        r, s = C.c_int32(), C.c_int32()
        self.lib.pcbgpu_get_tile(int(nl), C.byref(r), C.byref(s))
        return r.value, s.value

    @staticmethod
    def _tune_dict(t):
        return {k: getattr(t, k) for k, _ in Tune._fields_}

    def autotune(self, pairs, reps=2):
        """Picks tile rows, local sweeps and the delta band for the round [(grid, problem)] by timing it
        (pcbgpu_autotune) and installs them for its layer count. Returns the chosen setting."""
        qs = [p.query() for _, p in pairs]
        grids = (C.c_void_p * len(pairs))(*[g.h for g, _ in pairs])
        qp = (C.POINTER(Query) * len(pairs))(*[C.pointer(q) for q in qs])
        t = Tune()
        n = self.lib.pcbgpu_autotune(grids, qp, len(pairs), int(reps), C.byref(t))
        if n < 0:
            raise BackendError(f"autotune failed: {self.error()}")
        return self._tune_dict(t)

    def tune_get(self, nl):
        t = Tune()
        rc = self.lib.pcbgpu_tune_get(int(nl), C.byref(t))
        if rc < 0:
            raise BackendError("bad layer count")
        return self._tune_dict(t), bool(rc)

    def tune_set(self, nl, tile_rows=None, local_sweeps=4, delta=0):
        """Installs a setting for a layer count (tile_rows None: back to the defaults)."""
        if tile_rows is None:
            rc = self.lib.pcbgpu_tune_set(int(nl), None)
        else:
            t = Tune(int(tile_rows), int(local_sweeps), int(delta), 0, 0.0)
            rc = self.lib.pcbgpu_tune_set(int(nl), C.byref(t))
        if rc != 0:
            raise BackendError("bad tune setting")

    def tune_save(self, path):
        if self.lib.pcbgpu_tune_save(os.fsencode(path)) != 0:
            raise BackendError(self.error())

    def tune_load(self, path):
        n = self.lib.pcbgpu_tune_load(os.fsencode(path))
        if n < 0:
            raise BackendError(self.error())
        return n

    def _finish(self, out, rc, copy_field=True):
        if rc < 0 and rc != ERR_BUSY:
            err = self.error()
            self.lib.pcbgpu_out_release(C.byref(out))
            raise BackendError(f"solve failed ({rc}): {err}")
        res = Result(out, rc, copy_field)
        self.lib.pcbgpu_out_release(C.byref(out))
        return res

    def solve(self, prob, wait=True):
        out = Out()
        rc = self.lib.pcbgpu_solve(C.byref(prob.query()), C.byref(out), int(wait))
        return self._finish(out, rc)

    def grid(self, prob):
        return Grid(self, prob.cst, prob.st, prob.vcst, prob.vst)

    def solve_grid(self, grid, prob, wait=True):
        out = Out()
        rc = self.lib.pcbgpu_solve_grid(grid.h, C.byref(prob.query()), C.byref(out), int(wait))
        return self._finish(out, rc)

    def solve_rect_grid(self, grid, prob, rect=None, field=True):
        rx0, ry0, rw, rh = rect if rect is not None else (0, 0, prob.window[2], prob.window[3])
        out = Out()
        rc = self.lib.pcbgpu_solve_rect_grid(grid.h, C.byref(prob.query()), rx0, ry0, rw, rh, C.byref(out),
                                             0 if field else NOFIELD)
        return self._finish(out, rc, field)

    def round(self, pairs, field=True, rects=None):
        """Solves [(grid, problem)] as one round (pcbgpu_round_*): a list of Results in submission order. rects:
        one fixed rectangle per search (no widening), or None."""
        r = self.lib.pcbgpu_round_begin()
        if not r:
            raise BackendError(self.error())
        try:
            if not field:
                self.lib.pcbgpu_round_flags(r, NOFIELD)
            tickets = []
            for k, (g, p) in enumerate(pairs):
                if rects is not None:
                    rx0, ry0, rw, rh = rects[k] if rects[k] is not None else (0, 0, p.window[2], p.window[3])
                    t = self.lib.pcbgpu_round_submit_rect(r, g.h, C.byref(p.query()), rx0, ry0, rw, rh)
                else:
                    t = self.lib.pcbgpu_round_submit(r, g.h, C.byref(p.query()))
                if t < 0:
                    raise BackendError(self.error())
                tickets.append(t)
            if self.lib.pcbgpu_round_launch(r) != 0:
                raise BackendError(self.error())
            out = []
            for t in tickets:
                o = Out()
                rc = self.lib.pcbgpu_round_result(r, t, C.byref(o))
                out.append(self._finish(o, rc, field))
            return out
        finally:
            self.lib.pcbgpu_round_end(r)

    def backtrace(self, prob, res, via_gapc=0, via_gap2=0.0, old_vias=()):
        """Chain from res.best back to a start: list of (window state, code) or a negative error code."""
        out = Out()
        out.rx0, out.ry0, out.rw, out.rh = res.rect
        out.nl, out.ww, out.wh, out.best = res.nl, res.ww, res.wh, res.best
        d = np.ascontiguousarray(res.d, np.int32)
        out.d = d.ctypes.data_as(C.POINTER(C.c_int32))
        cap = max(1, d.size)
        states = np.zeros(cap, np.uint32)
        codes = np.zeros(cap, np.uint8)
        ov = np.asarray(list(old_vias), np.uint32)
        n = self.lib.pcbgpu_backtrace(C.byref(prob.query()), C.byref(out), int(via_gapc), float(via_gap2),
                                      ov.ctypes.data_as(C.POINTER(C.c_uint32)), len(ov),
                                      states.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      codes.ctypes.data_as(C.POINTER(C.c_uint8)), cap)
        if n < 0:
            return int(n)
        return [(int(states[k]), int(codes[k])) for k in range(n)]
