"""Pins the CPU oracle, which is the checker of the HIP back end.

Each check is chosen so that a plausible mistake in the oracle fails it:
  - closed forms on empty grids with a different weight for every move code (a swapped direction, a transposed
    operand or a diagonal priced like a straight step changes the numbers);
  - the cost of a move is the entered state's factor, not the left one's (a 1-row grid with one dear cell);
  - a via costs via_cost times the via factor and lands only where the landing rules allow;
  - scipy's Dijkstra (a library routine) on the explicit edge list of random cases;
  - the reference-order sweep solver (a second algorithm) against Dijkstra;
  - Bellman optimality of every finite state of the canonical field;
  - known answers of the cost resolution: own copper, another net's copper, the prior's bias and corridor;
  - the backtrace's first-strict-minimum rule, restated naively here on small cases, and two ties worked by hand;
  - the widening's rectangles, try counts and costs worked by hand on walls placed on either side of its bound.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import dijkstra

from pcbroute_amd.abi import BLOCK, DX, DY, FREE, INF, MAX_LAYERS, Problem, default_dirw

from cases import random_case
from known_cases import TIES, WIDENING


def empty(nl, ny, nx, cost=1):
    return (np.full((nl, ny, nx), cost, np.int32), np.zeros((nl, ny, nx), np.uint16),
            np.ones((ny, nx), np.int32), np.zeros((ny, nx), np.uint16))


def test_closed_form_empty_grid_every_quadrant(oracle):
    nx, ny = 31, 23
    dirw = np.zeros((MAX_LAYERS, 8), np.int32)
    dirw[0] = [10, 11, 13, 14, 17, 19, 23, 29]   # +x -x +y -y (+,+) (+,-) (-,+) (-,-)
    x0, y0 = 12, 9
    p = Problem(*empty(1, ny, nx), starts=[(y0 * nx + x0, 0)], targets=[0], via_cost=-1, dirw=dirw)
    d, best, opt = oracle.solve_rect(p)
    # the target (0, 0) is the best, so the canonical field keeps D < opt; compare where D < opt
    w = dirw[0]
    for j in range(ny):
        for i in range(nx):
            dx, dy = i - x0, j - y0
            ox = w[0] if dx > 0 else w[1]
            oy = w[2] if dy > 0 else w[3]
            dg = {(True, True): w[4], (True, False): w[5], (False, True): w[6], (False, False): w[7]}[(dx > 0, dy > 0)]
            ax, ay = abs(dx), abs(dy)
            m = min(ax, ay)
            # a diagonal is taken only where it is cheaper than the two straight steps it replaces
            want = m * min(dg, ox + oy) + (ax - m) * ox + (ay - m) * oy
            if want < opt:
                assert d[0, j, i] == want, (i, j)
            elif (j, i) == (0, 0):
                assert d[0, j, i] == want == opt
            else:
                assert d[0, j, i] == INF
    assert best == 0


def test_move_pays_the_entered_cell(oracle):
    cst, st, vcst, vst = empty(1, 1, 5)
    cst[0, 0, 2] = 100
    dirw = default_dirw(1, 12, 17)
    p = Problem(cst, st, vcst, vst, starts=[(0, 0)], targets=[4], via_cost=-1, dirw=dirw)
    d, best, opt = oracle.solve_rect(p)
    assert d[0, 0].tolist() == [0, 12, 1212, 1224, 1236]
    assert opt == 1236 and best == 4


def test_via_closed_form_and_landing(oracle):
    nx, ny = 9, 7
    cst, st, vcst, vst = empty(3, ny, nx)
    vcst[:] = 3
    # layer 1 at (4, 3) is another net's copper (negative map value, static owner 5): no track may enter it and no
    # via may land on it
    st[1, 3, 4] = 5
    cst[1, 3, 4] = -2
    p = Problem(cst, st, vcst, vst, starts=[(3 * nx + 4, 0)], targets=[(2 * ny + 3) * nx + 4], via_cost=40,
                dirw=default_dirw(3, 12, 17))
    d, best, opt = oracle.solve_rect(p)
    assert opt == 120                       # one via of 40 x factor 3 straight down to layer 2
    assert d[2, 3, 4] == 120
    assert d[1, 3, 4] == INF
    p2 = Problem(cst, st, vcst, vst, starts=[(3 * nx + 4, 0)], targets=[(1 * ny + 3) * nx + 4], via_cost=40,
                 dirw=default_dirw(3, 12, 17))
    d2, best2, opt2 = oracle.solve_rect(p2)
    assert best2 == -1 and opt2 == INF
    # the same cell with a non-negative map value is entered by track (the owner only matters for copper)
    cst[1, 3, 4] = 1
    p3 = Problem(cst, st, vcst, vst, starts=[(3 * nx + 4, 0)], targets=[(1 * ny + 3) * nx + 4], via_cost=40,
                 dirw=default_dirw(3, 12, 17))
    d3, best3, opt3 = oracle.solve_rect(p3)
    assert opt3 == 120 + 12 + 12            # via onto layer 1 next door (no landing on the cell itself), step back


def test_resolution_known_answers(oracle):
    nx, ny, nl = 6, 2, 2
    cst, st, vcst, vst = empty(nl, ny, nx, cost=4)
    me = 3
    st[0, 0, 1], cst[0, 0, 1] = me, -7          # own copper: 7
    st[0, 0, 2], cst[0, 0, 2] = 9, -7           # another net's copper: blocked, no via landing
    cst[0, 0, 3] = INF                          # blocked map value, free static: blocked, via may still land
    cst[0, 0, 4] = 0                            # a zero factor reads as 1
    vst[0, 5], vcst[0, 5] = me, -2              # own via copper
    vst[1, 5], vcst[1, 5] = 9, -2               # another net's via copper
    data = np.zeros((nl, 1, 2), np.uint8)
    data[1, 0, 0] = 10
    data[1, 0, 1] = 200
    # bias 8 for prior value 0 down to 0 at 252 and above; layer 0 has no prior layer (value 255: no bias)
    prior = dict(data=data, pcol=np.array([0, 0, 0, 1, 1, 1]), prow=np.array([0, 0]), play=[-1, 1],
                 mlbias=np.clip(8 - np.arange(256) // 28, 0, 8), mthr=50, corridor=True)
    tgt = (1 * ny + 0) * nx + 1                 # layer 1, (1, 0): in the corridor's forbidden part, but a target
    p = Problem(cst, st, vcst, vst, starts=[(0, 0)], targets=[tgt], me=me, via_cost=5, prior=prior,
                dirw=default_dirw(nl))
    cf, vok, tg, vcf = oracle.resolve(p)
    assert cf[0, 0, :].tolist() == [4, 7, 0, 0, 1, 4]
    assert vok[0, 0, :].tolist() == [1, 1, 0, 1, 1, 1]
    # layer 1: prior value 10 (< mthr 50) on columns 0-2 forbids them except the target (bias 8); 200 on 3-5 adds 1
    assert cf[1, 0, :].tolist() == [0, 4 + 8, 0, 4 + 1, 4 + 1, 4 + 1]
    assert vok[1, 0, :].tolist() == [0, 1, 0, 1, 1, 1]
    assert vcf[0, :].tolist() == [5, 5, 5, 5, 5, 10]
    assert vcf[1, 5] == INF


def _graph(oracle, p):
    """Explicit edge list of the rectangle (the whole window) from the oracle's resolved factors."""
    cf, vok, tg, vcf = oracle.resolve(p)
    nl, rh, rw = cf.shape
    N = nl * rh * rw
    src, dst, wt = [], [], []
    idx = np.arange(N).reshape(nl, rh, rw)
    for l in range(nl):
        for dm in range(8):
            dx, dy = DX[dm], DY[dm]
            # p at (i, j) -> q at (i + dx, j + dy)
            qi = slice(max(dx, 0), rw + min(dx, 0)), slice(max(dy, 0), rh + min(dy, 0))
            pi = slice(max(-dx, 0), rw + min(-dx, 0)), slice(max(-dy, 0), rh + min(-dy, 0))
            qs = idx[l][qi[1], qi[0]]
            ps = idx[l][pi[1], pi[0]]
            c = cf[l][qi[1], qi[0]].astype(np.int64)
            m = c > 0
            src.append(ps[m]); dst.append(qs[m]); wt.append(c[m] * int(p.dirw[l, dm]))
    if p.via_cost >= 0:
        for l in range(nl):
            for l2 in range(nl):
                if l2 == l:
                    continue
                m = (vok[l2] > 0) & (vcf < INF)
                src.append(idx[l][m]); dst.append(idx[l2][m]); wt.append(vcf[m].astype(np.int64))
    S = N  # a super source with edge g0 + 1 to each start (scipy drops explicit zeros)
    WN = rw * rh
    for s, g in p.starts:
        if g >= INF:
            continue
        src.append(np.array([S])); dst.append(np.array([s])); wt.append(np.array([max(g, 0) + 1]))
    src, dst, wt = map(np.concatenate, (src, dst, wt))
    G = coo_matrix((wt.astype(np.float64), (src, dst)), shape=(N + 1, N + 1)).tocsr()
    D = dijkstra(G, directed=True, indices=S)[:N] - 1
    D[~np.isfinite(D)] = INF
    return D.astype(np.int64).reshape(nl, rh, rw), tg


def _canon(D, tg):
    flat = D.ravel().copy()
    t = np.flatnonzero(tg.ravel())
    opt = INF
    best = -1
    if t.size:
        v = flat[t]
        opt = int(v.min())
        if opt < INF:
            best = int(t[np.flatnonzero(v == opt)[0]])
        else:
            opt = INF
    out = np.where(flat < opt, flat, INF)
    if best >= 0:
        out[best] = opt
    return out.reshape(D.shape), best, opt


@pytest.mark.parametrize("seed", range(12))
def test_dijkstra_matches_scipy(oracle, seed):
    p = random_case(1000 + seed, nx=40 + seed, ny=30, nl=1 + seed % 3)
    # whole window only (the rectangle the oracle solves below)
    p = Problem(p.cst, p.st, p.vcst, p.vst, starts=p.starts, targets=p.targets, window=p.window, me=p.me,
                via_cost=p.via_cost, dirw=p.dirw, lay=p.lay[:p.nl], own_cf=p.own_cf, own_vcf=p.own_vcf,
                prior=p.prior)
    D, tg = _graph(oracle, p)
    want, wbest, wopt = _canon(D, tg)
    d, best, opt = oracle.solve_rect(p)
    assert (best, opt) == (wbest, wopt)
    np.testing.assert_array_equal(d, want)


@pytest.mark.parametrize("seed", range(10))
def test_sweeps_match_dijkstra(oracle, seed):
    p = random_case(2000 + seed, nx=70, ny=50, nl=1 + seed % 4)
    d1, b1, o1 = oracle.solve_rect(p)
    d2, b2, o2, sweeps = oracle.sweep_rect(p)
    assert (b1, o1) == (b2, o2)
    np.testing.assert_array_equal(d1, d2)
    assert sweeps >= 1


@pytest.mark.parametrize("seed", range(6))
def test_bellman_optimality(oracle, seed):
    """Every state below opt is its start cost or the cheapest predecessor plus the move."""
    p = random_case(3000 + seed, nx=50, ny=40, nl=2 + seed % 2)
    cf, vok, tg, vcf = oracle.resolve(p)
    d, best, opt = oracle.solve_rect(p)
    nl, rh, rw = d.shape
    g0 = np.full(d.shape, INF, np.int64)
    for s, g in p.starts:
        l, c = divmod(s, rw * rh)
        j, i = divmod(c, rw)
        g0[l, j, i] = min(g0[l, j, i], max(g, 0))
    D = d.astype(np.int64)
    for l in range(nl):
        for j in range(rh):
            for i in range(rw):
                v = D[l, j, i]
                if v >= min(opt, INF) and not (l * rh * rw + j * rw + i == best):
                    continue
                cands = [g0[l, j, i]]
                if cf[l, j, i]:
                    for dm in range(8):
                        pi, pj = i - DX[dm], j - DY[dm]
                        if 0 <= pi < rw and 0 <= pj < rh and D[l, pj, pi] < INF:
                            cands.append(D[l, pj, pi] + int(p.dirw[l, dm]) * int(cf[l, j, i]))
                if p.via_cost >= 0 and vok[l, j, i] and vcf[j, i] < INF:
                    for l2 in range(nl):
                        if l2 != l and D[l2, j, i] < INF:
                            cands.append(D[l2, j, i] + int(vcf[j, i]))
                assert v == min(cands), (l, i, j)


@pytest.mark.parametrize("seed", range(6))
def test_astar_finds_the_optimum(oracle, seed):
    p = random_case(4000 + seed, nx=80, ny=60, nl=1 + seed % 3)
    d, best, opt = oracle.solve_rect(p)
    a, pops = oracle.astar(p)
    assert a == opt
    assert pops > 0


def _naive_backtrace(p, cf, vok, vcf, d, best):
    nl, rh, rw = d.shape
    g0 = {}
    for s, g in p.starts:
        g0[s] = min(g0.get(s, INF), max(g, 0))
    chain = []
    q = best
    while True:
        l, c = divmod(q, rw * rh)
        j, i = divmod(c, rw)
        if q in g0 and d[l, j, i] >= g0[q]:
            chain.append((q, 255))
            return chain
        opts = []
        if p.lay[l] and cf[l, j, i]:
            for dm in range(8):
                pi, pj = i - DX[dm], j - DY[dm]
                if 0 <= pi < rw and 0 <= pj < rh:
                    opts.append((int(d[l, pj, pi]) + int(p.dirw[l, dm]) * int(cf[l, j, i]), (l * rh + pj) * rw + pi, dm))
        if p.via_cost >= 0 and vok[l, j, i] and vcf[j, i] < INF:
            for l2 in range(nl):
                if l2 != l:
                    opts.append((int(d[l2, j, i]) + int(vcf[j, i]), (l2 * rh + j) * rw + i, 8 + l2))
        m = min(o[0] for o in opts)
        first = next(o for o in opts if o[0] == m)   # the first strict minimum in code order
        assert m == d[l, j, i]
        chain.append((q, first[2]))
        q = first[1]


# seeds of random_case(seed, 40, 32, 2 + seed % 2) whose targets are reachable (5002 and 5007 are not)
FIRST_MIN_SEEDS = (5000, 5001, 5003, 5004, 5005, 5006, 5008, 5009)


@pytest.mark.parametrize("seed", FIRST_MIN_SEEDS)
def test_backtrace_first_minimum(oracle, seed):
    p = random_case(seed, nx=40, ny=32, nl=2 + seed % 2, features=False)
    res = oracle.solve(p)
    assert res["best"] >= 0
    assert res["rect"] == (0, 0, p.window[2], p.window[3])
    cf, vok, tg, vcf = oracle.resolve(p)
    chain = oracle.backtrace(p, res["rect"], res["best"], res["d"])
    assert chain == _naive_backtrace(p, cf, vok, vcf, res["d"], res["best"])
    # the chain's moves add up to the target's cost
    s0 = chain[-1][0]
    g = min(g for s, g in p.starts if s == s0)
    assert res["d"].ravel()[res["best"]] >= g


def test_backtrace_via_spacing_rules(oracle):
    nx, ny = 12, 5
    cst, st, vcst, vst = empty(2, ny, nx)
    cst[0, :, 6] = INF                          # a wall on layer 0 at column 6: the path must via around it
    cst[1, :, 3] = INF
    cst[1, :, 9] = INF
    p = Problem(cst, st, vcst, vst, starts=[(2 * nx + 0, 0)], targets=[2 * nx + 11], via_cost=10,
                dirw=default_dirw(2))
    res = oracle.solve(p)
    ok = oracle.backtrace(p, res["rect"], res["best"], res["d"])
    vias = [c for _, c in ok if 8 <= c < 255]
    assert len(vias) == 2
    assert oracle.backtrace(p, res["rect"], res["best"], res["d"], via_gapc=20) == -2   # too close along the path
    # an old via of the net 1 cell away from one of the new ones
    via_cells = [s % (nx * ny) for s, c in ok if 8 <= c < 255]
    old = via_cells[0] + 1
    assert oracle.backtrace(p, res["rect"], res["best"], res["d"], via_gap2=4.0, old_vias=[old]) == -2
    assert oracle.backtrace(p, res["rect"], res["best"], res["d"], via_gap2=0.5, old_vias=[old]) == ok


def test_widening_keeps_the_optimum(oracle):
    """A near pair in a big window is solved on a small rectangle first; the result's cost equals the whole
    window's."""
    from pcbroute_amd import synth
    board = synth.kernel_board(11, nx=600, ny=500, nl=2)
    cst, st, vcst, vst = board
    rng = np.random.default_rng(3)
    for k in range(4):
        while True:
            i, j = int(rng.integers(100, 500)), int(rng.integers(100, 400))
            if st[0, j, i] == FREE and st[0, j + 20, i + 25] == FREE:
                break
        p = Problem(cst, st, vcst, vst, starts=[(j * 600 + i, 0)], targets=[(j + 20) * 600 + i + 25], via_cost=600)
        res = oracle.solve(p)
        d, best, opt = oracle.solve_rect(p)
        assert res["rect"][2] < 600 or res["tries"] > 1
        assert res["best"] == best
        rx0, ry0, rw, rh = res["rect"]
        l, c = divmod(best, 600 * 500)
        jj, ii = divmod(c, 600)
        assert res["d"][l, jj - ry0, ii - rx0] == opt


@pytest.mark.parametrize("case", range(len(WIDENING)))
def test_widening_known_rectangles(oracle, case):
    """The rectangles, try counts and costs worked out by hand from the widening rule (tests/known_cases.py)."""
    name, build, (ti, tj), rect, tries, cost = WIDENING[case]
    p = build()
    res = oracle.solve(p)
    assert res["rect"] == rect, name
    assert res["tries"] == tries, name
    assert res["best"] == tj * p.window[2] + ti, name
    assert res["d"][0, tj - rect[1], ti - rect[0]] == cost, name


@pytest.mark.parametrize("case", range(len(TIES)))
def test_backtrace_known_ties(oracle, case):
    """Ties between moves, and between a move and a via, broken as worked out by hand (tests/known_cases.py)."""
    name, build = TIES[case]
    p, chain = build()
    res = oracle.solve(p)
    assert res["best"] == chain[0][0]
    assert oracle.backtrace(p, res["rect"], res["best"], res["d"]) == chain, name
