"""HIP back end against the CPU oracle, bit for bit: canonical field, solved rectangle, best target, backtrace.

Every case prints its seed and shape before it touches the GPU."""
import numpy as np
import pytest

from pcbroute_amd import synth
from pcbroute_amd.abi import INF, Problem

from cases import random_case
from known_cases import TIES, WIDENING

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _same(res, ref):
    assert res.rect == ref["rect"], (res.rect, ref["rect"])
    assert res.best == ref["best"], (res.best, ref["best"])
    bad = np.argwhere(res.d != ref["d"])
    assert bad.size == 0, f"{len(bad)} states differ, first {bad[:5].tolist()}: gpu {res.d[tuple(bad[0])]} ref {ref['d'][tuple(bad[0])]}"


@pytest.mark.parametrize("seed", range(24))
def test_random_cases(backend, oracle, seed):
    nl = 1 + seed % 4
    p = random_case(seed, nx=96 + 8 * (seed % 5), ny=70 + 9 * (seed % 3), nl=nl)
    _say("case", seed, p.window, p.nl)
    ref = oracle.solve(p)
    res = backend.solve(p)
    if ref is None:
        pytest.fail("oracle could not solve the case")
    _same(res, ref)


@pytest.mark.parametrize("seed", range(16))
def test_random_cases_symmetric_weights(backend, oracle, seed):
    """The kernel variant for sign-symmetric move weights (what the router produces)."""
    p = random_case(900 + seed, nx=100 + 9 * (seed % 4), ny=74 + 7 * (seed % 3), nl=1 + seed % 5, sym=True)
    _say("sym case", seed, p.window, p.nl)
    _same(backend.solve(p), oracle.solve(p))


@pytest.mark.parametrize("rows,lsweeps", [(16, 1), (16, 2), (32, 8), (32, 64), (16, 64)])
def test_tile_shapes_do_not_change_results(backend, oracle, rows, lsweeps):
    p = random_case(100 + rows + lsweeps, nx=150, ny=131, nl=2)
    _say("tiles", rows, lsweeps)
    ref = oracle.solve(p)
    backend.set_tile(rows, lsweeps)
    try:
        res = backend.solve(p)
    finally:
        backend.set_tile(-1, -1)  # the defaults again
    _same(res, ref)


@pytest.mark.parametrize("delta", [1, 200, 3000, -1])
def test_delta_stepping_does_not_change_results(backend, oracle, delta):
    probs = [random_case(700 + delta % 97 + k, nx=150, ny=120, nl=1 + k % 3) for k in range(4)]
    _say("delta", delta)
    backend.set_delta(delta)
    try:
        for p in probs:
            _same(backend.solve(p), oracle.solve(p))
    finally:
        backend.set_delta(0)


@pytest.mark.parametrize("seed", range(6))
def test_grid_resident_solves(backend, oracle, seed):
    p = random_case(200 + seed, nx=120, ny=100, nl=1 + seed % 3)
    _say("grid", seed)
    g = backend.grid(p)
    try:
        res = backend.solve_grid(g, p)
        ref = oracle.solve(p)
        _same(res, ref)
        rect = (3, 2, p.window[2] - 5, p.window[3] - 4)
        rr = backend.solve_rect_grid(g, p, rect)
        d, best, opt = oracle.solve_rect(p, rect)
        assert rr.rect == rect
        np.testing.assert_array_equal(rr.d, d)
        if best >= 0:
            rw, rh = rect[2], rect[3]
            l, c = divmod(best, rw * rh)
            assert rr.best == (l * p.window[3] + c // rw + rect[1]) * p.window[2] + c % rw + rect[0]
        else:
            assert rr.best == -1
    finally:
        g.close()


# seeds of random_case(seed, 100, 90, 2 + seed % 2) whose targets are reachable (302, 303 and 313 are not)
BACKTRACE_SEEDS = (300, 301, 304, 305, 306, 307, 308, 309)


@pytest.mark.parametrize("seed", BACKTRACE_SEEDS)
def test_backtrace_matches_oracle(backend, oracle, seed):
    p = random_case(seed, nx=100, ny=90, nl=2 + seed % 2)
    _say("backtrace", seed)
    res = backend.solve(p)
    assert res.best >= 0
    gap, gap2 = seed % 3, float(seed % 4)
    mine = backend.backtrace(p, res, gap, gap2)
    ref = oracle.backtrace(p, res.rect, res.best, res.d, gap, gap2)
    assert mine == ref


@pytest.mark.parametrize("case", range(len(WIDENING)))
def test_widening_known_rectangles(backend, case):
    """The rectangles, try counts and costs worked out by hand from the widening rule (tests/known_cases.py)."""
    name, build, (ti, tj), rect, tries, cost = WIDENING[case]
    p = build()
    _say("widening", name)
    res = backend.solve(p)
    assert (res.rect, res.tries, res.best) == (rect, tries, tj * p.window[2] + ti), name
    assert res.d[0, tj - rect[1], ti - rect[0]] == cost, name


@pytest.mark.parametrize("case", range(len(TIES)))
def test_backtrace_known_ties(backend, case):
    name, build = TIES[case]
    p, chain = build()
    _say("tie", name)
    res = backend.solve(p)
    assert backend.backtrace(p, res) == chain, name


def test_unreachable_target(backend, oracle):
    p = random_case(400, nx=64, ny=64, nl=2, features=False)
    # wall the target in on every layer: its cell and ring blocked, no vias
    cst = p.cst.copy()
    st = p.st.copy()
    cst[:, 20:31, 20:31] = INF
    st[:, 20:31, 20:31] = 0xFFFF
    from pcbroute_amd.abi import Problem
    q = Problem(cst, st, p.vcst, p.vst, starts=[(p.state(0, 2, 2), 0)], targets=[p.state(1, 25, 25)], via_cost=-1)
    _say("unreachable")
    res = backend.solve(q)
    ref = oracle.solve(q)
    assert res.best == -1
    _same(res, ref)


def test_start_is_target(backend, oracle):
    p = random_case(401, nx=80, ny=60, nl=2, features=False)
    from pcbroute_amd.abi import Problem
    s = p.state(1, 40, 30)
    q = Problem(p.cst, p.st, p.vcst, p.vst, starts=[(s, 5)], targets=[s, p.state(0, 70, 50)])
    _say("start is target")
    res = backend.solve(q)
    ref = oracle.solve(q)
    _same(res, ref)
    assert res.best == s


@pytest.mark.parametrize("k", range(2))
def test_full_size_kernel_board(backend, oracle, k):
    """Two of the 64 searches of the 2000 x 1600 x 4 workload, whole grid, against the oracle's Dijkstra."""
    board = synth.kernel_board(1)
    p = synth.pair_problems(board, 7, 2)[k]
    _say("full size", k)
    g = backend.grid(p)
    try:
        res = backend.solve_rect_grid(g, p)
    finally:
        g.close()
    d, best, opt = oracle.solve_rect(p)
    assert res.best == best
    np.testing.assert_array_equal(res.d, d)


def test_grid_rect_uploads_batched(backend, oracle):
    """Rectangles of new maps uploaded in one call (pcbgpu_grid_upload_rects, a commit's footprint on several layers
    and the via grid) leave the grid holding the merged maps: its solves equal those of the merged maps uploaded
    whole, and the oracle's."""
    a = random_case(1300, nx=120, ny=90, nl=3)
    b = random_case(1301, nx=120, ny=90, nl=3)
    rects = [(0, 10, 5, 40, 30), (2, 61, 40, 50, 45), (3, 0, 0, 120, 7), (1, 119, 89, 1, 1)]  # layer 3: the vias
    cst, st, vcst, vst = a.cst.copy(), a.st.copy(), a.vcst.copy(), a.vst.copy()
    for l, i0, j0, w, h in rects:
        if l < 3:
            cst[l, j0:j0 + h, i0:i0 + w] = b.cst[l, j0:j0 + h, i0:i0 + w]
            st[l, j0:j0 + h, i0:i0 + w] = b.st[l, j0:j0 + h, i0:i0 + w]
        else:
            vcst[j0:j0 + h, i0:i0 + w] = b.vcst[j0:j0 + h, i0:i0 + w]
            vst[j0:j0 + h, i0:i0 + w] = b.vst[j0:j0 + h, i0:i0 + w]
    from pcbroute_amd.abi import Problem
    merged = Problem(cst, st, vcst, vst, starts=a.starts, targets=a.targets, window=a.window, me=a.me,
                     via_cost=a.via_cost, dirw=a.dirw, lay=a.lay[:a.nl], own_cf=a.own_cf, own_vcf=a.own_vcf)
    g, g2 = backend.grid(a), backend.grid(merged)
    try:
        g.upload_rects(rects, b.cst, b.st, b.vcst, b.vst)
        _say("batched rect uploads", len(rects))
        got, want = backend.solve_grid(g, merged), backend.solve_grid(g2, merged)
    finally:
        g.close()
        g2.close()
    assert got.rect == want.rect and got.best == want.best
    np.testing.assert_array_equal(got.d, want.d)
    _same(got, oracle.solve(merged))


def test_grid_uploads_outside_the_grid_are_refused(backend, oracle):
    """Uploads whose rectangle leaves the grid are refused on the host before any copy, also where the corner's
    coordinate sum passes 2^31 (it would wrap negative in 32 bits)."""
    import ctypes as C
    a = random_case(1302, nx=64, ny=40, nl=2)
    g = backend.grid(a)
    lib = backend.lib
    c = np.zeros(16, np.int32)
    s = np.zeros(16, np.uint16)
    cp, sp = c.ctypes.data_as(C.POINTER(C.c_int32)), s.ctypes.data_as(C.POINTER(C.c_uint16))
    big = 2**31 - 8
    try:
        for l, i0, j0, w, h in [(0, big, 0, 16, 1), (0, 0, big, 1, 16), (0, 60, 0, 5, 1), (3, 0, 0, 1, 1)]:
            _say("refused rect", (l, i0, j0, w, h))
            assert lib.pcbgpu_grid_upload_rect(g.h, l, i0, j0, w, h, cp, sp) == -1
        assert lib.pcbgpu_grid_upload(g.h, 0, big, 16, cp, sp) == -1
        rects = np.array([0, big, 0, 16, 1], np.int32)
        cps, sps = (C.POINTER(C.c_int32) * 1)(cp), (C.POINTER(C.c_uint16) * 1)(sp)
        assert lib.pcbgpu_grid_upload_rects(g.h, 1, rects.ctypes.data_as(C.POINTER(C.c_int32)), cps, sps) == -1
        got = backend.solve_grid(g, a)
    finally:
        g.close()
    _same(got, oracle.solve(a))


def test_window_tiles_equal_the_throughput_setting(backend, oracle):
    """A routing window solved with the window default (32 rows, 2 local sweeps) and with an explicit 16-row,
    4-sweep setting: the same rectangle, best state and field."""
    p = random_case(1400, nx=300, ny=220, nl=4, sym=True)
    auto = backend.solve(p)
    backend.set_tile(16, 4)
    try:
        fixed = backend.solve(p)
    finally:
        backend.set_tile(-1, -1)
    assert auto.rect == fixed.rect and auto.best == fixed.best
    np.testing.assert_array_equal(auto.d, fixed.d)
    _same(auto, oracle.solve(p))


@pytest.mark.parametrize("nl", [6, 7, 8])
def test_many_layers(backend, oracle, nl):
    """Six to eight layers (the C3 and C5 boards' counts), both kernel variants."""
    for k, sym in enumerate((False, True)):
        p = random_case(1200 + 10 * nl + k, nx=130, ny=97, nl=nl, sym=sym)
        _say("layers", nl, "sym", sym, p.window)
        _same(backend.solve(p), oracle.solve(p))


def test_c5_size_search(backend, oracle):
    """One search over a whole 3200 x 2400 x 8 grid (61.4M states, the size of the C5 board) against the oracle's
    Dijkstra, bit for bit."""
    board = synth.kernel_board(5, nx=3200, ny=2400, nl=8)
    p = synth.pair_problems(board, 11, 1)[0]
    _say("c5 size")
    g = backend.grid(p)
    try:
        res = backend.solve_rect_grid(g, p)
    finally:
        g.close()
    d, best, opt = oracle.solve_rect(p)
    assert res.best == best
    np.testing.assert_array_equal(res.d, d)


@pytest.mark.parametrize("nl", [2, 4, 6])
def test_round_equals_one_at_a_time(backend, oracle, nl):
    """A round of searches with different windows and two grids (two net classes) gives each search exactly its
    one-at-a-time result."""
    probs = [random_case(600 + 10 * nl + k, nx=140, ny=110, nl=nl) for k in range(10)]
    grids = [backend.grid(probs[0]), backend.grid(probs[1])]
    try:
        pairs = [(grids[k % 2], _on_grid(p, probs[k % 2])) for k, p in enumerate(probs)]
        _say("round", nl, len(pairs))
        got = backend.round(pairs)
        for (g, p), r in zip(pairs, got):
            one = backend.solve_grid(g, p)
            # pass counts may differ by a few (the field is relaxed in place, in whatever order workgroups visit
            # tiles); rectangle, best state and field may not
            assert r.rect == one.rect and r.best == one.best and r.sweeps > 0 and one.sweeps > 0
            np.testing.assert_array_equal(r.d, one.d)
            ref = oracle.solve(p)
            _same(r, ref)
    finally:
        for g in grids:
            g.close()


def test_round_search_without_target_fails_alone(backend, oracle):
    """A round scans its target bitmaps at launch: a search without a target fails at its own result (with the
    reason) while the whole-grid searches beside it (bitmaps large enough to be scanned for their nonzero words, on
    several host threads) still solve exactly."""
    import ctypes as C
    from pcbroute_amd.abi import Out
    board = synth.kernel_board(5, nx=1200, ny=1000, nl=4)
    good = synth.pair_problems(board, 9, 4)
    p0 = good[0]
    none = Problem(p0.cst, p0.st, p0.vcst, p0.vst, starts=p0.starts, targets=[], window=p0.window)
    g = backend.grid(p0)
    lib = backend.lib
    try:
        r = lib.pcbgpu_round_begin()
        qs = [good[0].query(), none.query()] + [p.query() for p in good[1:]]
        tickets = [lib.pcbgpu_round_submit(r, g.h, C.byref(q)) for q in qs]
        assert tickets == [0, 1, 2, 3, 4]
        assert lib.pcbgpu_round_launch(r) == 0
        rcs, errs = [], []
        for t in tickets:
            o = Out()
            rcs.append(lib.pcbgpu_round_result(r, t, C.byref(o)))
            errs.append(backend.error())
            lib.pcbgpu_out_release(C.byref(o))
        lib.pcbgpu_round_end(r)
        assert rcs[1] < 0 and all(rc > 0 for k, rc in enumerate(rcs) if k != 1)
        assert "no start or no target" in errs[1]
        for p, got in zip(good, backend.round([(g, q) for q in good], field=True)):
            _same(got, oracle.solve(p))
    finally:
        g.close()


def test_c3_round_against_the_oracle(backend, oracle):
    """The batched speculative round at the C3 board's shape, with fewer and smaller searches: a 2400 x 1800 x 6 board
    seen by two net classes (clearances of 1 and 2 cells), 32 searches of 2-8 pins whose boxes are log-uniform in
    64..512 cells plus a 40-cell margin, tuned by the autotuner on another draw and then run as one round (passes
    replayed from captured hipGraph chunks). Every search's rectangle, best state and field equal the oracle's
    Dijkstra with widening, and the first four equal their lone solves."""
    grids = synth.class_grids(3)
    searches = synth.net_searches(grids, 17, 32, smax=512)
    sample = synth.net_searches(grids, 1017, 12, smax=512)
    gs = [backend.grid(Problem(c, s, vc, vs, starts=[], targets=[])) for c, s, vc, vs in grids]
    nl = grids[0][0].shape[0]
    try:
        _say("c3 autotune", len(sample))
        t = backend.autotune([(gs[k], p) for k, p in sample], reps=1)
        assert backend.get_tile(nl) == (t["tile_rows"], t["local_sweeps"])
        pairs = [(gs[k], p) for k, p in searches]
        _say("c3 round", len(pairs), "setting", t["tile_rows"], t["local_sweeps"], t["delta"])
        got = backend.round(pairs)
        assert sum(r.best >= 0 for r in got) >= len(pairs) // 2
        for k, ((g, p), r) in enumerate(zip(pairs, got)):
            _say("c3 search", k, p.window)
            _same(r, oracle.solve(p))
            if k < 4:  # a lone solve gives the same result
                one = backend.solve_grid(g, p)
                assert (one.rect, one.best) == (r.rect, r.best)
                np.testing.assert_array_equal(one.d, r.d)
    finally:
        backend.tune_set(nl, None)
        for g in gs:
            g.close()


def _on_grid(p, other):
    """p's search on other's maps (another net class's grid of the same board size)."""
    from pcbroute_amd.abi import Problem
    return Problem(other.cst, other.st, other.vcst, other.vst, starts=p.starts, targets=p.targets, window=p.window,
                   me=p.me, via_cost=p.via_cost, dirw=p.dirw, lay=p.lay[:p.nl], own_cf=p.own_cf, own_vcf=p.own_vcf,
                   prior=p.prior)


def test_widening_solves_match(backend, oracle):
    """Near pairs in a big window: the first rectangle is small and may widen; rectangle, best and field match."""
    from pcbroute_amd.abi import Problem
    board = synth.kernel_board(11, nx=600, ny=500, nl=2)
    cst, st, vcst, vst = board
    rng = np.random.default_rng(3)
    probs = []
    for k in range(4):
        while True:
            i, j = int(rng.integers(100, 500)), int(rng.integers(100, 400))
            if st[0, j, i] == 0 and st[0, j + 20, i + 25] == 0:
                break
        probs.append(Problem(cst, st, vcst, vst, starts=[(j * 600 + i, 0)], targets=[(j + 20) * 600 + i + 25]))
    g = backend.grid(probs[0])
    try:
        for p in probs:
            _say("widening")
            ref = oracle.solve(p)
            _same(backend.solve(p), ref)
            _same(backend.solve_grid(g, p), ref)
        got = backend.round([(g, p) for p in probs])
        for p, r in zip(probs, got):
            _same(r, oracle.solve(p))
    finally:
        g.close()


def test_cpp_gpu_h_shim(tmp_path):
    """The C++ gpu.h shim (GpuQuery / GpuOut / gpu_solve) gives what the C ABI gives: tests/shim_check.cpp, built here
    with g++ against the library."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib = os.path.join(root, "pcbroute_amd")
    exe = str(tmp_path / "shim_check")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(root, "tests", "shim_check.cpp"), f"-L{lib}",
                        "-lpcbgpu", f"-Wl,-rpath,{lib}"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _say("cpp shim")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


def test_autotune_picks_a_setting_and_keeps_results(backend, oracle):
    """The autotuner times a sample round under its candidates, installs the fastest for the layer count, and the
    round's results stay those of the oracle under the picked setting."""
    nl = 3
    probs = [random_case(1300 + k, nx=160, ny=120, nl=nl, sym=k % 2 == 0) for k in range(6)]
    g = backend.grid(probs[0])
    try:
        pairs = [(g, _on_grid(p, probs[0])) for p in probs]
        _say("autotune", len(pairs))
        t = backend.autotune(pairs, reps=1)
        assert t["tried"] >= 2 and t["ms"] > 0
        assert t["tile_rows"] in (16, 32) and 1 <= t["local_sweeps"] <= 8 and t["delta"] >= -1
        got, tuned = backend.tune_get(nl)
        assert tuned and (got["tile_rows"], got["local_sweeps"], got["delta"]) == (t["tile_rows"], t["local_sweeps"], t["delta"])
        assert backend.get_tile(nl) == (t["tile_rows"], t["local_sweeps"])
        for (_, p), r in zip(pairs, backend.round(pairs)):
            _same(r, oracle.solve(p))
    finally:
        backend.tune_set(nl, None)
        g.close()
    assert backend.tune_get(nl)[1] is False
