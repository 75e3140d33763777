"""The C ABI: libpcbgpu.so loads and exports every function the headers under include/ declare, and the ctypes mirror
of the structs has the C layout (offsets printed by a C program compiled here). No device calls."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from pcbroute_amd import LIB_PATH, load_library
from pcbroute_amd.abi import Entry, Out, Query

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pcbgpu.h")
HEADERS = sorted(os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))


def declared():
    names = set()
    for h in HEADERS:
        names |= set(re.findall(r"\b((?:pcbgpu|pcbprior)_[a-z0-9_]+)\s*\(", open(h).read()))
    return sorted(names)


def test_header_declares_the_reference_entry_points():
    names = declared()
    for n in ("pcbgpu_available", "pcbgpu_configure", "pcbgpu_solve", "pcbgpu_out_release", "pcbgpu_out_at",
              "pcbgpu_backtrace"):
        assert n in names


def test_library_exports_every_declared_function():
    if not os.path.exists(LIB_PATH):
        pytest.fail("libpcbgpu.so is not built (python -m pcbroute_amd.build)")
    lib = load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT ((?:pcbgpu|pcbprior)_[a-z0-9_]+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    for n in declared():
        assert getattr(lib, n) is not None
    assert "pcbprior_train_step" in declared()


def test_struct_layout_matches_c():
    from pcbroute_amd.router import Options, Result
    classes = {"pcbgpu_query": Query, "pcbgpu_out": Out, "pcbgpu_entry": Entry, "pcbr_options": Options,
               "pcbr_result": Result}
    fields = {s: [f[0] for f in cls._fields_] for s, cls in classes.items()}
    router_h = os.path.join(os.path.dirname(HEADER), "pcbroute.h")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', f'#include "{router_h}"',
             'int main(void) {']
    for s, fs in fields.items():
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f in fs:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-o", exe, src], check=True)
        got = dict(line.rsplit(" ", 1) for line in subprocess.run([exe], capture_output=True, text=True,
                                                                   check=True).stdout.splitlines())
    for s, cls in classes.items():
        assert int(got[f"{s} size"]) == C.sizeof(cls), s
        for f in fields[s]:
            assert int(got[f"{s}.{f}"]) == getattr(cls, f).offset, (s, f)


def test_tune_table_set_save_load(tmp_path):
    """The autotuner's table (host only: no device call): install, read back, save, clear, load."""
    from pcbroute_amd import Backend
    b = Backend()
    try:
        b.tune_set(4, 16, 3, -1)
        t, tuned = b.tune_get(4)
        assert tuned and (t["tile_rows"], t["local_sweeps"], t["delta"]) == (16, 3, -1)
        f = tmp_path / "tune.txt"
        b.tune_save(str(f))
        assert "4 16 3 -1" in f.read_text()
        b.tune_set(4, None)
        assert b.tune_get(4)[1] is False
        assert b.tune_load(str(f)) == 1
        t, tuned = b.tune_get(4)
        assert tuned and (t["tile_rows"], t["local_sweeps"], t["delta"]) == (16, 3, -1)
        with pytest.raises(Exception):
            b.tune_set(4, 24, 3, 0)  # rows must be 16 or 32
        with pytest.raises(Exception):
            b.tune_set(4, 16, 0, 0)  # at least one local sweep
        with pytest.raises(Exception):
            b.tune_get(9)
    finally:
        b.tune_set(4, None)


def test_tile_setting_and_defaults():
    """A tile setting (host only) reads back, (-1, -1) restores the defaults, bad settings are refused."""
    from pcbroute_amd import Backend
    b = Backend()
    default = b.get_tile(4)
    try:
        b.set_tile(32, 3)
        assert b.get_tile(4) == (32, 3)
        b.set_tile(0, 5)  # 0 keeps the rows
        assert b.get_tile(4) == (32, 5)
        with pytest.raises(Exception):
            b.set_tile(24, 2)
        with pytest.raises(Exception):
            b.set_tile(16, -3)
    finally:
        b.set_tile(-1, -1)
    assert b.get_tile(4) == default and default[1] == 2


def test_out_at_reads_the_solved_rectangle_only():
    """pcbgpu_out_at (host only) maps a window state into the solved rectangle and answers INF outside it, also for
    state indices past the last layer whose layer number would not fit 32 bits (GpuOut::operator(), src/gpu.h:39-50)."""
    import numpy as np
    lib = load_library()
    INF = (1 << 30) - 1
    nl, ww, wh, rx0, ry0, rw, rh = 2, 10, 7, 3, 2, 4, 3
    d = np.arange(nl * rh * rw, dtype=np.int32)
    o = Out(rx0=rx0, ry0=ry0, rw=rw, rh=rh, nl=nl, ww=ww, wh=wh, best=-1,
            d=d.ctypes.data_as(C.POINTER(C.c_int32)), cap=d.nbytes)
    for l in range(nl):
        for j in range(wh):
            for i in range(ww):
                q = (l * wh + j) * ww + i
                inside = rx0 <= i < rx0 + rw and ry0 <= j < ry0 + rh
                want = int(d[(l * rh + j - ry0) * rw + i - rx0]) if inside else INF
                assert lib.pcbgpu_out_at(C.byref(o), q) == want
    WN = ww * wh
    for q in (-1, nl * WN, (1 << 32) * WN + (ry0 * ww + rx0), (1 << 62)):
        assert lib.pcbgpu_out_at(C.byref(o), q) == INF
