"""Hand-derived known answers for the rectangle widening and for backtrace ties.

Every expected value below is worked out in the comments from the rules themselves, not taken from a run:

Widening (src/gpu.cu:575-619): try k solves the box around the starts and targets grown by 80 * 3^k cells, clipped to
the window (x1, y1 exclusive); a first rectangle with 2 * area > window area is the whole window; a try is final when
its rectangle is the window or when every cell just outside it has a bound above the best cost, the bound being the
cheapest empty-grid walk from the starts' box to the cell plus from the cell to the targets' box at the least move
weights. The weights here are 12 straight, 17 diagonal, every cell factor 1, one layer, no vias.

Backtrace (src/router.cpp:1602-1679): from the best target, each step takes the first predecessor in code order (moves
+x -x +y -y (+,+) (+,-) (-,+) (-,-), then vias by the layer they come from) whose field value plus the step's cost
equals the state's own value; a start whose value is at least its g0 ends the chain with code 255.
"""
import numpy as np

from pcbroute_amd.abi import INF, MAX_LAYERS, Problem, default_dirw


def _empty(nl, ny, nx):
    return (np.full((nl, ny, nx), 1, np.int32), np.zeros((nl, ny, nx), np.uint16),
            np.ones((ny, nx), np.int32), np.zeros((ny, nx), np.uint16))


def _one_layer(nx, ny, start, target, edit=None):
    cst, st, vcst, vst = _empty(1, ny, nx)
    if edit is not None:
        edit(cst)
    return Problem(cst, st, vcst, vst, starts=[(start[1] * nx + start[0], 0)], targets=[target[1] * nx + target[0]],
                   via_cost=-1, dirw=default_dirw(1))


def _wall_gap_top(c):
    c[0, 21:181, 105] = INF         # column 105 blocked on rows 21..180; row 20 open


def _wall_gap_top_dear(c):
    _wall_gap_top(c)
    c[0, 20, 105] = 4               # the gap cell at factor 4


def _wall_whole_rect(c):
    c[0, 20:181, 105] = INF         # column 105 blocked on every row of the first rectangle


# (name, problem builder, target (i, j), expected rect (x0, y0, w, h), expected tries, expected cost at the target)
WIDENING = [
    # start (100, 100), target (110, 105) in 600 x 400: box x 100..110, y 100..105; try 0 grows by 80 to x 20..190,
    # y 20..185 -> (20, 20, 171, 166); 2 * 171 * 166 = 56772 <= 240000, so not the whole window. The cost is the octile
    # walk 5 diagonal + 5 straight = 5 * 17 + 5 * 12 = 145; the nearest outside cell (row 19 or column 19) is 80+ steps
    # from both boxes (>= 1920), so try 0 is final.
    ("near pair", lambda: _one_layer(600, 400, (100, 100), (110, 105)), (110, 105), (20, 20, 171, 166), 1, 145),
    # start (100, 100), target (110, 100), wall on column 105 rows 21..180, gap at row 20. Try 0: x 20..190, y 20..180
    # -> (20, 20, 171, 161). The path runs through the gap (105, 20): each half is 5 diagonal + 75 straight = 985, so 1970.
    # The cheapest outside cell is on row 19 or 181: 81 rows from both boxes, |i - 100| + |i - 110| >= 10 columns, the
    # bound 2 * 81 * 12 + 5 * 10 = 1994 > 1970 (each diagonal step trades a straight 12 for 17), so try 0 is final.
    ("gap inside, final", lambda: _one_layer(600, 400, (100, 100), (110, 100), _wall_gap_top), (110, 100),
     (20, 20, 171, 161), 1, 1970),
    # The same with the gap cell at factor 4: through it the path costs at least 980 + 4 * 12 + 985 = 2013 > 1994, so try
    # 0 is not final; try 1 grows by 240: x 0..350, y 0..340 -> (0, 0, 351, 341). Going round row 19 or row 181 costs
    # 2 * (5 * 17 + 76 * 12) = 1994, below 2013; the outside cells of try 1 are 240+ steps away, so it is final.
    ("gap too dear", lambda: _one_layer(600, 400, (100, 100), (110, 100), _wall_gap_top_dear), (110, 100),
     (0, 0, 351, 341), 2, 1994),
    # The wall covers every row of try 0: no target there, so widen; try 1 as above, round row 19 or 181: 1994.
    ("no path in first", lambda: _one_layer(600, 400, (100, 100), (110, 100), _wall_whole_rect), (110, 100),
     (0, 0, 351, 341), 2, 1994),
    # start (50, 50), target (60, 60) in 200 x 150: try 0 is x 0..140, y 0..140 (141 x 141 = 19881), 2 * 19881 = 39762 >
    # 30000: the whole window. Cost 10 diagonal = 170.
    ("over half", lambda: _one_layer(200, 150, (50, 50), (60, 60)), (60, 60), (0, 0, 200, 150), 1, 170),
    # The first pair in 228 x 249: 2 * 171 * 166 = 56772 = 228 * 249 exactly, which is not over half.
    ("exactly half", lambda: _one_layer(228, 249, (100, 100), (110, 105)), (110, 105), (20, 20, 171, 166), 1, 145),
    # One row less, 228 * 248 = 56544 < 56772: over half, the whole window.
    ("just over half", lambda: _one_layer(228, 248, (100, 100), (110, 105)), (110, 105), (0, 0, 228, 248), 1, 145),
]


def tie_square():
    """3 x 3, one layer, straight 12 and diagonal 24 (= two straight), start (0, 0), target (2, 2): every monotone
    path costs 48. At (2, 2) the +x move from (1, 2) (36 + 12), the +y move from (2, 1) and the diagonal from (1, 1)
    (24 + 24) tie: the first, +x, wins. At (1, 2): +x from (0, 2) (24 + 12), +y from (1, 1), diagonal from (0, 1)
    (12 + 24) tie: +x again. (0, 2) and (0, 1) have only +y (their other neighbours cost more), ending at the start.
    Chain (state, code): (8, 0) (7, 0) (6, 2) (3, 2) (0, 255)."""
    cst, st, vcst, vst = _empty(1, 3, 3)
    dirw = np.zeros((MAX_LAYERS, 8), np.int32)
    dirw[0, :4] = 12
    dirw[0, 4:] = 24
    p = Problem(cst, st, vcst, vst, starts=[(0, 0)], targets=[8], via_cost=-1, dirw=dirw)
    return p, [(8, 0), (7, 0), (6, 2), (3, 2), (0, 255)]


def tie_via():
    """2 x 1, two layers, start layer 0 cell (0, 0) (state 0), target layer 1 cell (1, 0) (state 3), straight 12, vias
    cost V = 10. At the target the +x move from layer 1 (0, 0) (V + 12) and the via from layer 0 (1, 0) (12 + V) tie:
    moves come before vias, so +x. At layer 1 (0, 0) (state 2) only the via from layer 0 (code 8 + 0) gives V.
    Chain: (3, 0) (2, 8) (0, 255)."""
    cst, st, vcst, vst = _empty(2, 1, 2)
    p = Problem(cst, st, vcst, vst, starts=[(0, 0)], targets=[3], via_cost=10, dirw=default_dirw(2))
    return p, [(3, 0), (2, 8), (0, 255)]


TIES = [("square", tie_square), ("via", tie_via)]
