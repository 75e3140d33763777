"""The router's host logic on the CPU: seeded board generation, and whole routes over the oracle's exact search
(oracle/libroute_oracle.so) against the committed fixture; waves with read-block validation against one net at a
time; the independent check of the routes."""
import json
import os

import numpy as np
import pytest

import oracle
from pcbroute_amd import router as R

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "route_golden.json")


@pytest.fixture(scope="module")
def lib():
    return oracle.build_route_oracle()


def test_board_generation_is_seeded(lib):
    a = R.Board(R.spec("c4", seed=5), lib_path=lib)
    b = R.Board(R.spec("c4", seed=5), lib_path=lib)
    c = R.Board(R.spec("c4", seed=6), lib_path=lib)
    assert a.info.hash == b.info.hash != c.info.hash
    i = a.info
    assert (i.nx, i.ny, i.nl) == (1200, 1000, 4)
    pads = a.pads()
    assert len(pads) == i.n_pads and i.n_parts >= 60
    assert np.all(pads[:, 0] >= 0) and np.all(pads[:, 2] <= i.nx) and np.all(pads[:, 1] >= 0) and np.all(pads[:, 3] <= i.ny)
    assert np.all(pads[:, 2] > pads[:, 0]) and np.all(pads[:, 3] > pads[:, 1])
    nets = pads[:, 5]
    counts = np.bincount(nets[nets >= 0])
    assert len(counts) == i.n_nets and counts.min() >= 2 and counts.max() <= 6 and counts.sum() == i.n_pins
    assert i.n_connections == i.n_pins - i.n_nets
    # pads do not overlap
    m = np.zeros((i.ny, i.nx), np.int32)
    for x0, y0, x1, y1, _, _ in pads:
        m[y0:y1, x0:x1] += 1
    assert m.max() == 1


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tiny_boards_route_as_the_fixture(lib, seed):
    g = [x for x in json.load(open(GOLDEN)) if x["kind"] == "tiny" and x["seed"] == seed][0]
    b = R.Board(R.spec("tiny", seed=seed), lib_path=lib)
    assert b.info.hash == g["board_hash"]
    r = R.Router(b, R.default_options(R.library(lib)), lib_path=lib)
    res = r.route()
    for k in ("route_hash", "opens", "violations", "searches", "length_cells", "vias", "iterations"):
        assert res[k] == g[k], k
    assert r.eval() == (0, 0)


@pytest.mark.parametrize("wave", [2, 5])
def test_waves_equal_one_net_at_a_time(lib, wave):
    b = R.Board(R.spec("tiny", seed=2), lib_path=lib)
    seq = R.Router(b, R.default_options(R.library(lib)), lib_path=lib).route()
    wav = R.Router(b, R.default_options(R.library(lib), wave=wave), lib_path=lib).route()
    assert wav["route_hash"] == seq["route_hash"]
    assert wav["revalidated"] > 0  # the nets of this board interact: some searches were redone
    assert wav["rounds"] < wav["searches"]


def test_routes_are_connected_paths(lib):
    b = R.Board(R.spec("tiny", seed=3), lib_path=lib)
    r = R.Router(b, R.default_options(R.library(lib)), lib_path=lib)
    r.route()
    pads = b.pads()
    for n in range(b.info.n_nets):
        rows = r.net_route(n)
        paths = np.split(rows, np.where(rows[:, 0] < 0)[0])
        for p in paths:
            p = p[p[:, 0] >= 0]
            if len(p) < 2:
                continue
            d = np.abs(np.diff(p, axis=0))
            step = (d[:, 0] == 0) & (np.maximum(d[:, 1], d[:, 2]) == 1)   # a track step on one layer
            via = (d[:, 0] > 0) & (d[:, 1] == 0) & (d[:, 2] == 0)          # a via
            assert np.all(step | via)
            # each path ends on a pad of its net
            l, i, j = p[-1]
            on = [(x0 <= i < x1 and y0 <= j < y1 and (pl < 0 or pl == l)) for x0, y0, x1, y1, pl, pn in pads if pn == n]
            assert any(on)


def test_group_waves_equal_one_at_a_time(lib):
    b = R.Board(R.spec("tiny", seed=2), lib_path=lib)
    seq = R.Router(b, R.default_options(R.library(lib), group_mm=3.0), lib_path=lib).route()
    grp = R.Router(b, R.default_options(R.library(lib), group_mm=3.0, wave=-8), lib_path=lib).route()
    assert grp["route_hash"] == seq["route_hash"]


@pytest.mark.parametrize("negotiate", [1, 2])
def test_parallel_negotiation_commits_whole_batches(lib, negotiate):
    """negotiate 1 / 2: a batch's nets are all routed against the board as the batch began and all committed (more
    searches per GPU round than one net at a time); overlaps are left to the next round and legalizing leaves the
    board clean, as the independent check confirms."""
    b = R.Board(R.spec("tiny", seed=2), lib_path=lib)
    seq = R.Router(b, R.default_options(R.library(lib)), lib_path=lib).route()
    r = R.Router(b, R.default_options(R.library(lib), wave=64, negotiate=negotiate), lib_path=lib)
    par = r.route()
    assert par["violations"] == 0 and r.eval() == (par["opens"], 0)
    assert par["searches"] / par["rounds"] > 2 * seq["searches"] / seq["rounds"]
    if negotiate == 2:
        assert par["revalidated"] == 0  # nothing is validated, nothing redone


def test_training_samples(lib):
    b = R.Board(R.spec("tiny", seed=1), lib_path=lib)
    r = R.Router(b, R.default_options(R.library(lib)), lib_path=lib)
    r.record_samples(32, 64)
    r.route()
    x, t, m = r.samples()
    nl = b.info.nl
    assert x.shape == (b.info.n_nets, 32, 64, R.FEAT_PER_LAYER * nl + R.FEAT_NET) and t.shape == (b.info.n_nets, 32, 64, nl)
    assert x.min() >= 0 and x.max() <= 1 and set(np.unique(t)) <= {0.0, 1.0}
    assert np.all(m.sum(axis=(1, 2)) > 0)
    for k in range(len(x)):
        own_pads = x[k, :, :, 3::R.FEAT_PER_LAYER][..., :nl].max(axis=-1)  # this net's pad cells, any layer
        assert own_pads.max() > 0
        mst = x[k, :, :, R.FEAT_PER_LAYER * nl]
        assert mst[own_pads > 0].max() > 0.8                      # the spanning tree passes through the pins
        assert t[k].sum() > 0 and (t[k].max(axis=-1) * mst).sum() > 0  # the route lies along it
