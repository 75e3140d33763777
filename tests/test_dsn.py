"""DSN in, SES out (src/dsn.cpp:75-384, write_ses src/router.cpp:2938-2963) on the CPU: boards read from DSN text
(placement, rotation, back side, through-hole pads, classes, keepouts, a polygon outline) checked against positions
worked out by hand; generated boards written as DSN and read back unchanged; routes written as sessions and judged by
the reference's own evaluator (`pcbroute --eval`, built from its sources into oracle/_ref/ by oracle/build_ref.sh;
those checks skip where it cannot be built)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from pcbroute_amd import router as R

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "pcbroute")

# 20 x 10 mm, two layers, one net from one end to the other (the board of the reference's tests/test_gpu_gate.py)
GATE = """(pcb test.dsn
  (resolution um 10)
  (unit um)
  (structure
    (layer F.Cu (type signal))
    (layer B.Cu (type signal))
    (boundary (path pcb 0  0 0  20000 0  20000 10000  0 10000  0 0))
    (via V)
    (rule (width 200) (clearance 200))
  )
  (placement
    (component P (place A1 2000 2000 front 0) (place B1 18000 8000 front 0))
  )
  (library
    (image P (pin SMD 1 0 0))
    (padstack SMD (shape (rect F.Cu -300 -300 300 300)) (attach off))
    (padstack V (shape (circle F.Cu 600)) (shape (circle B.Cu 600)) (attach off))
  )
  (network
    (net N1 (pins A1-1 B1-1))
    (class default N1 (circuit (use_via V)) (rule (width 200) (clearance 200)))
  )
)
"""

# an L-shaped 2-layer board: SMD resistors at 0 and 90 degrees, one on the back and a wider one at 45 degrees (a pad
# turned off the axes is its bounding rectangle here: the wide pitch keeps the two rectangles apart), two through-hole headers
# (one turned 180 degrees, a pin of its own rotated), two classes with their own vias, a keepout on the top layer
FEATURES = """(pcb feat.dsn
  (parser (string_quote ")
    (space_in_quoted_tokens on))
  (resolution um 10)
  (unit um)
  (structure
    (layer F.Cu (type signal))
    (layer B.Cu (type signal))
    (boundary (path pcb 0  0 0  30000 0  30000 12000  16000 12000  16000 20000  0 20000  0 0))
    (via V1 V2)
    (rule (width 200) (clearance 200))
    (keepout "" (rect F.Cu 8000 4000 10000 16000))
  )
  (placement
    (component R (place R1 3000 3000 front 0) (place R2 25000 6000 front 90) (place R3 4000 17000 back 0))
    (component W (place R4 12000 2500 front 45))
    (component T (place J1 3000 10000 front 0) (place J2 13000 17000 front 180))
  )
  (library
    (image R (pin SMDR 1 -800 0) (pin SMDR 2 800 0))
    (image W (pin SMDR 1 -1500 0) (pin SMDR 2 1500 0))
    (image T (pin TH 1 0 0) (pin TH (rotate 90) 2 0 2540))
    (padstack SMDR (shape (rect F.Cu -400 -500 400 500)) (attach off))
    (padstack TH (shape (circle F.Cu 1600)) (shape (circle B.Cu 1600)) (attach off))
    (padstack V1 (shape (circle F.Cu 600)) (shape (circle B.Cu 600)) (attach off))
    (padstack V2 (shape (circle F.Cu 800)) (shape (circle B.Cu 800)) (attach off))
  )
  (network
    (net GND (pins R1-1 R2-1 J1-1))
    (net "Net-(R1-Pad2)" (pins R1-2 R3-1 J2-2))
    (net N3 (pins R2-2 R4-1 J2-1))
    (net N4 (pins R3-2 R4-2 J1-2))
    (class default N3 N4 (circuit (use_via V1)) (rule (width 200) (clearance 200)))
    (class power GND "Net-(R1-Pad2)" (circuit (use_via V2)) (rule (width 400) (clearance 250)))
  )
)
"""


@pytest.fixture(scope="module")
def lib():
    return oracle.build_route_oracle()


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF) and os.path.isdir("/root/reference/src"):
        subprocess.run(["bash", os.path.join(os.path.dirname(REF), "..", "build_ref.sh")], check=True, capture_output=True)
    if not os.path.exists(REF):
        pytest.skip("the reference's evaluator is not built (its sources are absent)")
    return REF


def _ref_eval(ref, dsn, ses):
    out = subprocess.run([ref, "-de", dsn, "--eval", ses, "-v"], capture_output=True, text=True, timeout=300)
    m = re.search(r"eval: (\d+) nets, (\d+) connections, (\d+) open, (\d+) violations", out.stdout)
    assert m, out.stdout + out.stderr
    return tuple(int(x) for x in m.groups()), out.stderr


def _write(tmp_path, name, text):
    p = str(tmp_path / name)
    open(p, "w").write(text)
    return p


def test_gate_board_reads_and_routes_clean(lib, tmp_path, ref):
    d = _write(tmp_path, "gate.dsn", GATE)
    b = R.Board.from_dsn(d, cell_um=50, lib_path=lib)
    i = b.info
    assert (i.nx, i.ny, i.nl, i.n_pads, i.n_nets, i.n_connections) == (400, 200, 2, 2, 1, 1)
    # pad A1: copper 1700..2300 um = cells 34..46; cells whose centres lie inside: 34..45
    assert b.pads().tolist() == [[34, 34, 46, 46, 0, 0], [354, 154, 366, 166, 0, 0]]
    r = R.Router(b, R.default_options(R.library(lib)), lib_path=lib)
    res = r.route()
    assert res["opens"] == 0 and res["violations"] == 0
    s = str(tmp_path / "gate.ses")
    r.write_ses(s)
    (nets, conns, opens, viol), log = _ref_eval(ref, d, s)
    assert (nets, conns, opens, viol) == (1, 1, 0, 0), log


def test_placement_rotation_side_and_classes(lib, tmp_path):
    b = R.Board.from_dsn(_write(tmp_path, "feat.dsn", FEATURES), cell_um=50, lib_path=lib)
    i = b.info
    assert (i.nx, i.ny, i.nl, i.n_pads, i.n_nets, i.n_pins) == (600, 400, 2, 12, 4, 12)
    pads = b.pads()

    def cells(x0, y0, x1, y1):  # the cells whose centres lie inside a copper rectangle (um)
        return [int(np.ceil(x0 / 50 - 0.5)), int(np.ceil(y0 / 50 - 0.5)), int(np.floor(x1 / 50 - 0.5)) + 1,
                int(np.floor(y1 / 50 - 0.5)) + 1]

    # R1 at 0 degrees: pin 1 at (2200, 3000), 800 x 1000 um, top layer, net GND (0)
    assert pads[0].tolist() == cells(1800, 2500, 2600, 3500) + [0, 0]
    # R2 at 90 degrees: pin 1 (-800, 0) turns to (0, -800), the pad to 1000 x 800
    assert pads[2].tolist() == cells(24500, 4800, 25500, 5600) + [0, 0]
    # R3 on the back: x mirrored (pin 1 at +800), the pad on the bottom layer, net "Net-(R1-Pad2)" (1)
    assert pads[4].tolist() == cells(4400, 16500, 5200, 17500) + [1, 1]
    # J1 through-hole: every layer; a circle's cells are those of its inscribed square
    q = 800 / np.sqrt(2)
    assert pads[8].tolist() == cells(3000 - q, 10000 - q, 3000 + q, 10000 + q) + [-1, 0]
    # J2 turned 180 degrees: pin 2 (0, 2540) lands at (13000, 17000 - 2540)
    assert pads[11].tolist() == cells(13000 - q, 14460 - q, 13000 + q, 14460 + q) + [-1, 1]


def test_feature_board_routes_clean_by_the_reference(lib, tmp_path, ref):
    """Routes on the feature board avoid the keepout and stay inside the L outline, by the reference's own judge."""
    d = _write(tmp_path, "feat.dsn", FEATURES)
    b = R.Board.from_dsn(d, cell_um=50, lib_path=lib)
    r = R.Router(b, R.default_options(R.library(lib)), lib_path=lib)
    res = r.route()
    assert res["violations"] == 0 and r.eval() == (res["opens"], 0)
    s = str(tmp_path / "feat.ses")
    r.write_ses(s)
    assert '(net "Net-(R1-Pad2)"' in open(s).read()  # the name as the DSN wrote it
    (nets, conns, opens, viol), log = _ref_eval(ref, d, s)
    assert (nets, conns) == (4, 8) and viol == 0 and opens == res["opens"], log


@pytest.mark.parametrize("kind,seed", [("tiny", 2), ("c1", 1)])
def test_generated_boards_through_dsn(lib, tmp_path, ref, kind, seed):
    """A generated board written as DSN reads back with the same pads and nets; its routes, written as a session,
    are clean by the reference's judge with the connection count both sides agree on."""
    b = R.Board(R.spec(kind, seed=seed), lib_path=lib)
    d = str(tmp_path / f"{kind}.dsn")
    b.write_dsn(d)
    b2 = R.Board.from_dsn(d, cell_um=R.spec(kind, seed=seed).cell_um, lib_path=lib)
    assert np.array_equal(b.pads(), b2.pads())
    i, i2 = b.info, b2.info
    assert (i.nx, i.ny, i.nl, i.n_nets, i.n_pins, i.n_connections) == (i2.nx, i2.ny, i2.nl, i2.n_nets, i2.n_pins,
                                                                       i2.n_connections)
    if kind == "c1":
        return  # the route of the C1 board over the oracle takes long; its session is judged in the GPU tests
    r = R.Router(b, R.default_options(R.library(lib)), lib_path=lib)
    res = r.route()
    s = str(tmp_path / f"{kind}.ses")
    r.write_ses(s)
    (nets, conns, opens, viol), log = _ref_eval(ref, d, s)
    assert conns == i.n_connections and opens == res["opens"] and viol == 0, log


@pytest.mark.parametrize("board", ["feat", "tiny"])
def test_smoothed_session_is_shorter_and_clean(lib, tmp_path, ref, board):
    """The smoothed session (the default, as the reference's smooth_routes, src/smooth.cpp:105-225) holds fewer wires
    and no more track than the grid paths' straight runs of the same routes, and the reference's judge finds it as
    connected and without violations."""
    if board == "feat":
        d = _write(tmp_path, "feat.dsn", FEATURES)
        b = R.Board.from_dsn(d, cell_um=50, lib_path=lib)
    else:
        b = R.Board(R.spec("tiny", seed=2), lib_path=lib)
        d = str(tmp_path / "tiny.dsn")
        b.write_dsn(d)
    got = {}
    for smooth in (0, 1):
        r = R.Router(b, R.default_options(R.library(lib), smooth=smooth), lib_path=lib)
        res = r.route()
        s = str(tmp_path / f"s{smooth}.ses")
        r.write_ses(s)
        out = subprocess.run([ref, "-de", d, "--eval", s], capture_output=True, text=True, timeout=300).stdout
        m = re.search(r"eval: \d+ nets, (\d+) connections, (\d+) open, (\d+) violations, ([\d.]+) mm of track, (\d+) vias",
                      out)
        assert m, out
        got[smooth] = dict(route=res["route_hash"], opens=res["opens"], conns=int(m.group(1)), open=int(m.group(2)),
                           viol=int(m.group(3)), track=float(m.group(4)), vias=int(m.group(5)),
                           wires=open(s).read().count("(wire "))
    raw, sm = got[0], got[1]
    assert raw["route"] == sm["route"]  # smoothing only rewrites the session
    assert sm["viol"] == 0 and raw["viol"] == 0
    assert sm["open"] == raw["open"] == raw["opens"] and sm["conns"] == raw["conns"]
    assert sm["vias"] == raw["vias"]
    assert sm["wires"] < raw["wires"] and sm["track"] <= raw["track"] + 1e-6


def test_dsn_errors(lib, tmp_path):
    cases = {"unbalanced": GATE[:-3], "no_structure": "(pcb x (placement))", "unit": GATE.replace("(unit um)", "(unit parsec)"),
             "no_via": GATE.replace("(via V)", ""), "no_image": GATE.replace("(image P", "(image Q")}
    for name, text in cases.items():
        with pytest.raises(R.BackendError):
            R.Board.from_dsn(_write(tmp_path, name + ".dsn", text), cell_um=50, lib_path=lib)
    with pytest.raises(R.BackendError):
        R.Board.from_dsn(str(tmp_path / "missing.dsn"), lib_path=lib)
    with pytest.raises(R.BackendError):
        R.Board.from_dsn(_write(tmp_path, "gate.dsn", GATE), cell_um=0, lib_path=lib)


# the board of the reference's tests/test_image_keepouts.py: two through-hole pins and a part whose image is a hole
# (a keepout circle 3 mm across on both layers) between them
HOLE = """(pcb test.dsn
  (resolution um 10)
  (unit um)
  (structure
    (layer F.Cu (type signal))
    (layer B.Cu (type signal))
    (boundary (path pcb 0  0 0  20000 0  20000 10000  0 10000  0 0))
    (via V)
    (rule (width 200) (clearance 200))
  )
  (placement
    (component P (place A1 2000 5000 front 0) (place B1 18000 5000 front 0))
    (component M (place H1 10000 5000 front 0))
  )
  (library
    (image P (pin TH 1 0 0))
    (image M (keepout "" (circle F.Cu 3000)) (keepout "" (circle B.Cu 3000)))
    (padstack TH (shape (circle F.Cu 800)) (shape (circle B.Cu 800)) (attach off))
    (padstack V (shape (circle F.Cu 600)) (shape (circle B.Cu 600)) (attach off))
  )
  (network
    (net N1 (pins A1-1 B1-1))
    (class default N1 (circuit (use_via V)) (rule (width 200) (clearance 200)))
  )
)
"""


def test_route_goes_around_an_image_keepout(lib, tmp_path, ref):
    """A part's image keepout (a hole) is placed with the part and obeyed: the straight line between the pins crosses
    it, the route goes around, and the reference's judge sees no keepout violation."""
    d = _write(tmp_path, "hole.dsn", HOLE)
    b = R.Board.from_dsn(d, cell_um=50, lib_path=lib)
    r = R.Router(b, R.default_options(R.library(lib)), lib_path=lib)
    res = r.route()
    assert res["opens"] == 0 and res["violations"] == 0
    rows = r.net_route(0)
    rows = rows[rows[:, 0] >= 0]
    near = np.hypot(rows[:, 1] + 0.5 - 200, rows[:, 2] + 0.5 - 100)  # cells from the hole's centre
    assert near.min() >= 30 + 2  # the hole's radius (1500 um) plus half the track width
    s = str(tmp_path / "hole.ses")
    r.write_ses(s)
    (nets, conns, opens, viol), log = _ref_eval(ref, d, s)
    assert (nets, conns, opens, viol) == (1, 1, 0, 0), log
