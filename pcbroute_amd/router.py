"""ctypes binding of the PathFinder router (include/pcbroute.h) that drives the search back end.

`Board` wraps a seeded synthetic board, `Router` routes it; both bind libpcbgpu.so unless a library path is given.
"""
import ctypes as C

import numpy as np

from . import LIB_PATH, BackendError
from .abi import Out, Query

MAX_CLASSES = 4
FEAT_PER_LAYER = 5
FEAT_NET = 2
PIX = 4


class BoardSpec(C.Structure):
    _fields_ = [("width_mm", C.c_double), ("height_mm", C.c_double), ("cell_um", C.c_double), ("layers", C.c_int32),
                ("n_soic16", C.c_int32), ("n_0805", C.c_int32), ("n_qfp64", C.c_int32), ("n_nets", C.c_int32),
                ("n_classes", C.c_int32), ("track_mm", C.c_double * MAX_CLASSES),
                ("clear_mm", C.c_double * MAX_CLASSES), ("via_mm", C.c_double * MAX_CLASSES),
                ("edge_mm", C.c_double), ("seed", C.c_uint64), ("n_bga256", C.c_int32), ("n_planes", C.c_int32)]


class BoardInfo(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("nx", "ny", "nl", "n_pads", "n_nets", "n_pins", "n_connections",
                                         "n_parts")] + [("hash", C.c_uint64)]


class Options(C.Structure):
    _fields_ = [("max_iters", C.c_int32), ("patience", C.c_int32), ("pres0", C.c_double),
                ("pres_growth", C.c_double), ("hist_inc", C.c_double), ("window_mm", C.c_double),
                ("via_cost_mm", C.c_double), ("wrong_way", C.c_double), ("wave", C.c_int32), ("group_mm", C.c_double),
                ("ml_weight", C.c_double), ("ml_sat", C.c_double), ("ml_corridor", C.c_double),
                ("ml_margin_mm", C.c_double), ("legalize", C.c_int32), ("verbose", C.c_int32),
                ("negotiate", C.c_int32), ("rescue", C.c_int32), ("smooth", C.c_int32),
                ("ml_rounds", C.c_int32), ("gpu_pops", C.c_int64), ("max_pops", C.c_int64),
                ("probe_pops", C.c_int64), ("gpu_wait", C.c_int32), ("time_limit_s", C.c_double),
                ("rescue_s", C.c_double), ("ripups", C.c_int32)]


class Result(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("nets", "connections", "opens", "violations", "conflict_nets",
                                         "iterations")] + \
               [(k, C.c_int64) for k in ("searches", "search_fallbacks", "rounds", "revalidated", "length_cells",
                                         "vias")] + \
               [("tile_visits", C.c_uint64)] + \
               [(k, C.c_double) for k in ("route_s", "search_s", "prior_s", "upload_s", "backtrace_s")] + \
               [("host_s", C.c_double * 4), ("route_hash", C.c_uint64)] + \
               [(k, C.c_int64) for k in ("cpu_searches", "gpu_searches", "gpu_failures", "chain_rejects",
                                         "probe_opens", "budget_opens")] + \
               [("cpu_pops", C.c_uint64), ("cpu_s", C.c_double), ("first_conflict_nets", C.c_int32),
                ("first_conflict_cells", C.c_int32), ("rescue_tried", C.c_int32), ("rescue_kept", C.c_int32),
                ("opens_before_rescue", C.c_int32), ("limited", C.c_int32), ("ripups", C.c_int32),
                ("solved_states", C.c_uint64)]


SEARCH_CB = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(Query), C.POINTER(Out))


def _bind(lib):
    lib.pcbr_board_generate.argtypes = [C.POINTER(BoardSpec)]
    lib.pcbr_board_generate.restype = C.c_void_p
    lib.pcbr_board_free.argtypes = [C.c_void_p]
    lib.pcbr_board_read_dsn.argtypes = [C.c_char_p, C.c_double]
    lib.pcbr_board_read_dsn.restype = C.c_void_p
    lib.pcbr_board_write_dsn.argtypes = [C.c_void_p, C.c_char_p]
    lib.pcbr_write_ses.argtypes = [C.c_void_p, C.c_char_p]
    lib.pcbr_board_get_info.argtypes = [C.c_void_p, C.POINTER(BoardInfo)]
    lib.pcbr_board_pads.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_size_t]
    lib.pcbr_default_options.argtypes = [C.POINTER(Options)]
    lib.pcbr_create.argtypes = [C.c_void_p, C.POINTER(Options)]
    lib.pcbr_create.restype = C.c_void_p
    lib.pcbr_destroy.argtypes = [C.c_void_p]
    lib.pcbr_route.argtypes = [C.c_void_p, C.POINTER(Result)]
    lib.pcbr_eval.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.pcbr_net_route.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.c_size_t]
    lib.pcbr_net_route.restype = C.c_int64
    lib.pcbr_set_search_callback.argtypes = [C.c_void_p, SEARCH_CB, C.c_void_p]
    lib.pcbr_set_prior.argtypes = [C.c_void_p, C.c_void_p]
    lib.pcbr_write_prior.argtypes = [C.c_void_p, C.c_char_p]
    lib.pcbr_read_prior.argtypes = [C.c_void_p, C.c_char_p]
    lib.pcbr_record_samples.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.pcbr_num_samples.argtypes = [C.c_void_p]
    lib.pcbr_num_samples.restype = C.c_int32
    F = C.POINTER(C.c_float)
    lib.pcbr_get_sample.argtypes = [C.c_void_p, C.c_int32, F, F, F]
    lib.pcbr_last_error.restype = C.c_char_p
    return lib


_libs = {}


def library(path=LIB_PATH):
    if path not in _libs:
        from . import load_library
        _libs[path] = _bind(load_library(path) if path == LIB_PATH else C.CDLL(path))
    return _libs[path]


def spec(kind="c4", seed=1, **kw):
    """Board specs of the configurations: 'c1' a 2-layer 40x30 mm board (6 SOIC-16, 20 0805, 60 nets), 'c4' a
    4-layer 60x50 mm board (2 QFP-64, 8 SOIC-16, 100 or more 0805 so that every one of its 150-300 nets has two pads), 'tiny' a 2-layer
    18x14 mm board for CPU checks, 'c5' the 8-layer 160x120 mm board (two reserved inner planes, four BGA-256, 10
    QFP-64, 30 SOIC-16, 400 0805, 1200 nets, classes of 0.1 / 0.15 / 0.25 mm track), 'c5small' the same mix on 50x40 mm
    (one BGA-256, 160 nets). All on a 50 um grid; one class of 0.2 mm track and clearance with 0.6 mm vias unless
    stated."""
    s = BoardSpec()
    s.cell_um, s.n_classes, s.edge_mm, s.seed = 50.0, 1, 0.5, seed
    s.track_mm[0], s.clear_mm[0], s.via_mm[0] = 0.2, 0.2, 0.6
    if kind == "c1":
        s.width_mm, s.height_mm, s.layers = 40.0, 30.0, 2
        s.n_soic16, s.n_0805, s.n_qfp64, s.n_nets = 6, 20, 0, 60
    elif kind == "c4":
        s.width_mm, s.height_mm, s.layers = 60.0, 50.0, 4
        s.n_nets = int(np.random.default_rng(seed).integers(150, 301))
        # two pads for every net at least: more 0805s on boards with more nets
        s.n_soic16, s.n_qfp64 = 8, 2
        s.n_0805 = max(100, (2 * s.n_nets + 20 - 8 * 16 - 2 * 64 + 1) // 2)
    elif kind in ("c5", "c5small"):
        # 8 layers, the middle two reserved as planes, BGA-256 parts that need fine-pitch escapes (their nets take
        # the 0.1 mm class), QFP-64 and SOIC-16 parts, 0805 passives, three net classes
        s.layers, s.n_planes, s.n_classes = 8, 2, 3
        for k, (t, c, v) in enumerate(((0.1, 0.1, 0.35), (0.15, 0.15, 0.5), (0.25, 0.2, 0.6))):
            s.track_mm[k], s.clear_mm[k], s.via_mm[k] = t, c, v
        if kind == "c5":
            s.width_mm, s.height_mm = 160.0, 120.0
            s.n_bga256, s.n_qfp64, s.n_soic16, s.n_0805, s.n_nets = 4, 10, 30, 400, 1200
        else:
            s.width_mm, s.height_mm = 50.0, 40.0
            s.n_bga256, s.n_qfp64, s.n_soic16, s.n_0805, s.n_nets = 1, 2, 4, 40, 160
    elif kind == "tiny":
        s.width_mm, s.height_mm, s.layers = 18.0, 14.0, 2
        s.n_soic16, s.n_0805, s.n_qfp64, s.n_nets = 1, 10, 0, 12
    else:
        raise ValueError(kind)
    for k, v in kw.items():
        if k in ("track_mm", "clear_mm", "via_mm"):
            for i, x in enumerate(v):
                getattr(s, k)[i] = x
        else:
            setattr(s, k, v)
    return s


def default_options(lib=None, **kw):
    o = Options()
    (lib or library()).pcbr_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Board:
    def __init__(self, spec_, lib_path=LIB_PATH, dsn=None, cell_um=50.0):
        """A seeded synthetic board of spec_, or (spec_ None) the board of the DSN file dsn on cell_um cells."""
        self.lib = library(lib_path)
        self.spec = spec_
        if spec_ is None:
            self.h = self.lib.pcbr_board_read_dsn(str(dsn).encode(), float(cell_um))
        else:
            self.h = self.lib.pcbr_board_generate(C.byref(spec_))
        if not self.h:
            raise BackendError(self.lib.pcbr_last_error().decode())
        self.info = BoardInfo()
        self.lib.pcbr_board_get_info(self.h, C.byref(self.info))

    @classmethod
    def from_dsn(cls, path, cell_um=50.0, lib_path=LIB_PATH):
        return cls(None, lib_path=lib_path, dsn=path, cell_um=cell_um)

    def write_dsn(self, path):
        if self.lib.pcbr_board_write_dsn(self.h, str(path).encode()) != 0:
            raise BackendError(self.lib.pcbr_last_error().decode())

    def pads(self):
        n = self.lib.pcbr_board_pads(self.h, None, 0)
        a = np.zeros((n, 6), np.int32)
        self.lib.pcbr_board_pads(self.h, a.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return a

    def close(self):
        if self.h:
            self.lib.pcbr_board_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Router:
    def __init__(self, board, options=None, lib_path=LIB_PATH):
        self.lib = library(lib_path)
        self.board = board
        self.options = options if options is not None else default_options(self.lib)
        self.h = self.lib.pcbr_create(board.h, C.byref(self.options))
        if not self.h:
            raise BackendError(self.lib.pcbr_last_error().decode())
        self._cb = None
        self._prior = None

    def close(self):
        if self.h:
            self.lib.pcbr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def on_search(self, fn):
        """fn(query pointer, out pointer) after every search (None: off)."""
        self._cb = SEARCH_CB(lambda user, q, o: fn(q, o)) if fn else SEARCH_CB()
        self.lib.pcbr_set_search_callback(self.h, self._cb, None)

    def set_prior(self, net):
        """Attach a pcbroute_amd.prior.PriorNet (in_ch = 5 x layers + 2, out_ch = layers), or None."""
        self._prior = net
        self.lib.pcbr_set_prior(self.h, net.h if net is not None else None)

    def write_ses(self, path):
        """The routes as a Specctra session file (after route)."""
        if self.lib.pcbr_write_ses(self.h, str(path).encode()) != 0:
            raise BackendError(self.lib.pcbr_last_error().decode())

    def write_prior(self, path):
        """The priors of the last route's first round as an MLP1 file; returns the number of nets written."""
        n = self.lib.pcbr_write_prior(self.h, str(path).encode())
        if n < 0:
            raise BackendError(self.lib.pcbr_last_error().decode())
        return n

    def read_prior(self, path):
        """Load an MLP1 prior file (used in the first round while no network is attached); returns the number of
        the board's nets it has a map for."""
        n = self.lib.pcbr_read_prior(self.h, str(path).encode())
        if n < 0:
            raise BackendError(self.lib.pcbr_last_error().decode())
        return n

    def record_samples(self, ch, cw):
        if self.lib.pcbr_record_samples(self.h, ch, cw) != 0:
            raise BackendError(self.lib.pcbr_last_error().decode())
        self._crop = (ch, cw)

    def route(self):
        r = Result()
        if self.lib.pcbr_route(self.h, C.byref(r)) != 0:
            raise BackendError(self.lib.pcbr_last_error().decode())
        out = {k: getattr(r, k) for k, _ in Result._fields_}
        out["host_s"] = dict(zip(("rebuild", "stamp", "reads", "query"), list(r.host_s)))
        return out

    def eval(self):
        o, v = C.c_int32(), C.c_int32()
        self.lib.pcbr_eval(self.h, C.byref(o), C.byref(v))
        return o.value, v.value

    def net_route(self, net):
        n = self.lib.pcbr_net_route(self.h, net, None, 0)
        a = np.zeros((max(n, 0), 3), np.int32)
        if n > 0:
            self.lib.pcbr_net_route(self.h, net, a.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return a

    def samples(self):
        """The recorded training samples: x (n, ch, cw, 5 nl + 2), target (n, ch, cw, nl), mask (n, ch, cw)."""
        n = self.lib.pcbr_num_samples(self.h)
        ch, cw = self._crop
        nl = self.board.info.nl
        x = np.zeros((n, ch, cw, FEAT_PER_LAYER * nl + FEAT_NET), np.float32)
        t = np.zeros((n, ch, cw, nl), np.float32)
        m = np.zeros((n, ch, cw), np.float32)
        F = C.POINTER(C.c_float)
        for k in range(n):
            self.lib.pcbr_get_sample(self.h, k, x[k].ctypes.data_as(F), t[k].ctypes.data_as(F), m[k].ctypes.data_as(F))
        return x, t, m
