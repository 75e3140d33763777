"""The CPU search (pcbgpu_cpu_search, astar.cpp) against the oracle: on seeded cases with every query feature switched
on at random, its best target and cost equal the oracle's canonical ones and its chain is the oracle's canonical
backtrace over the oracle's field, state for state, under both heuristics; budgets, unreachable targets, the via
spacing and bad queries. Host code only: no device."""
import numpy as np
import pytest

from pcbroute_amd import cpu_search
from pcbroute_amd.abi import ERR_BUDGET, ERR_CHAIN, ERR_FAIL, INF, Problem

from cases import random_case


def _oracle_answer(oracle, p, gap=0, gap2=0.0, old=()):
    ref = oracle.solve(p)
    if ref["best"] < 0:
        return ref, None
    d = ref["d"]
    rx0, ry0, rw, rh = ref["rect"]
    l, c = divmod(ref["best"], p.window[2] * p.window[3])
    j, i = divmod(c, p.window[2])
    opt = int(d[l, j - ry0, i - rx0])
    chain = oracle.backtrace(p, ref["rect"], ref["best"], d, gap, gap2, old)
    return ref, (opt, chain)


@pytest.mark.parametrize("heuristic", [0, 1])
@pytest.mark.parametrize("seed", range(20))
def test_cpu_search_equals_the_canonical_result(oracle, seed, heuristic):
    p = random_case(2000 + seed, nx=90 + 7 * (seed % 4), ny=70 + 5 * (seed % 3), nl=1 + seed % 4, sym=seed % 2 == 1)
    ref, ans = _oracle_answer(oracle, p)
    got = cpu_search(p, heuristic=heuristic)
    if ans is None:
        assert got["rc"] == 0 and got["best"] == -1
        return
    opt, chain = ans
    assert got["best"] == ref["best"] and got["opt"] == opt
    assert got["rc"] == len(chain) and got["chain"] == chain


@pytest.mark.parametrize("seed", range(8))
def test_pad_boxes_heuristic_and_a_tighter_search(oracle, seed):
    """Boxes around each target (what the router passes) give the same answer as one box per layer, with no more
    pops than a search without an estimate (boxes on every layer of the whole window make h = 0)."""
    p = random_case(2100 + seed, nx=120, ny=90, nl=2 + seed % 2, features=False)
    ww, wh = p.window[2], p.window[3]
    boxes = []
    for t in p.targets:
        l, c = divmod(t, ww * wh)
        j, i = divmod(c, ww)
        boxes.append((l, i, j, i, j))
    plain = cpu_search(p, boxes=[(-1, 0, 0, ww - 1, wh - 1)])
    tight = cpu_search(p, boxes=boxes)
    auto = cpu_search(p)
    for r in (tight, auto):
        assert (r["rc"], r["best"], r["opt"], r["chain"]) == (plain["rc"], plain["best"], plain["opt"], plain["chain"])
    assert tight["pops"] <= plain["pops"]


def test_budget_and_resume_answers(oracle):
    p = random_case(2200, nx=150, ny=120, nl=2, features=False)
    full = cpu_search(p)
    assert full["rc"] > 0 and full["pops"] > 20
    short = cpu_search(p, max_pops=10)
    assert short["rc"] == ERR_BUDGET and short["pops"] == 10
    assert cpu_search(p, max_pops=full["pops"])["rc"] == full["rc"]


def test_unreachable_target_is_exhausted(oracle):
    p = random_case(400, nx=64, ny=64, nl=2, features=False)
    cst, st = p.cst.copy(), p.st.copy()
    cst[:, 20:31, 20:31] = INF
    st[:, 20:31, 20:31] = 0xFFFF
    q = Problem(cst, st, p.vcst, p.vst, starts=[(p.state(0, 2, 2), 0)], targets=[p.state(1, 25, 25)], via_cost=-1)
    assert oracle.solve(q)["best"] == -1
    r = cpu_search(q)
    assert r["rc"] == 0 and r["best"] == -1


def test_start_on_a_target(oracle):
    p = random_case(401, nx=80, ny=60, nl=2, features=False)
    s = p.state(1, 40, 30)
    q = Problem(p.cst, p.st, p.vcst, p.vst, starts=[(s, 5)], targets=[s, p.state(0, 70, 50)])
    r = cpu_search(q)
    assert r["best"] == s and r["opt"] == 5 and r["chain"] == [(s, 255)]


@pytest.mark.parametrize("seed", range(6))
def test_via_spacing_rejects_the_same_chains(oracle, seed):
    """With via spacing rules the CPU search's chain is checked exactly as the backtrace checks the device's."""
    p = random_case(2300 + seed, nx=100, ny=80, nl=3, features=False)
    gap, gap2 = 4 + seed % 3, float(9 + seed)
    ref, ans = _oracle_answer(oracle, p, gap, gap2, old=[5 * p.nx + 7, 40 * p.nx + 50])
    got = cpu_search(p, via_gapc=gap, via_gap2=gap2, old_vias=[5 * p.nx + 7, 40 * p.nx + 50])
    if ans is None:
        assert got["rc"] == 0
        return
    opt, chain = ans
    if chain == ERR_CHAIN:
        assert got["rc"] == ERR_CHAIN
    else:
        assert got["chain"] == chain


def test_bad_queries_are_refused():
    p = random_case(2400, nx=40, ny=30, nl=2, features=False)
    q = p.query()
    q.dirw[0][0] = 0
    from pcbroute_amd import load_library
    import ctypes as C
    lib = load_library()
    best, opt, pops = C.c_int64(), C.c_int32(), C.c_uint64()
    states = np.zeros(8, np.uint32)
    codes = np.zeros(8, np.uint8)
    rc = lib.pcbgpu_cpu_search(C.byref(q), 0, None, 0, 1, 0, 0, 0.0, None, 0, C.byref(best), C.byref(opt), C.byref(pops),
                               states.ctypes.data_as(C.POINTER(C.c_uint32)), codes.ctypes.data_as(C.POINTER(C.c_uint8)),
                               8)
    assert rc == ERR_FAIL
    assert cpu_search(p, heuristic=7)["rc"] == ERR_FAIL


@pytest.mark.parametrize("seed", range(4))
def test_least_factor_scales_the_estimate(oracle, seed):
    """Cost maps whose factors are all at least 8 (as the router's are): with cf_min 8 the same answer as with 1 in
    fewer pops; a cf_min above a factor the search meets fails instead of answering."""
    p = random_case(2500 + seed, nx=140, ny=100, nl=2 + seed % 2, features=False)
    cst = np.where(p.cst < INF, p.cst * 8, p.cst).astype(np.int32)
    vcst = np.where(p.vcst < INF, p.vcst * 8, p.vcst).astype(np.int32)
    q = Problem(cst, p.st, vcst, p.vst, starts=p.starts, targets=p.targets, window=p.window, via_cost=p.via_cost,
                dirw=p.dirw, lay=p.lay[:p.nl])
    ref, ans = _oracle_answer(oracle, q)
    one, eight = cpu_search(q, cf_min=1), cpu_search(q, cf_min=8)
    assert (one["rc"], one["best"], one["opt"], one["chain"]) == (eight["rc"], eight["best"], eight["opt"], eight["chain"])
    if ans is not None:
        assert eight["opt"] == ans[0] and eight["chain"] == ans[1]
        assert eight["pops"] < one["pops"]
        assert cpu_search(q, cf_min=9)["rc"] == ERR_FAIL
