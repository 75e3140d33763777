"""Seeded synthetic search inputs for the single-search kernel workload.

The board: an nl-layer nx x ny grid with int32 cell costs of base 1; a fraction of each layer's cells blocked as
random 4-connected blobs of 3-30 cells; congestion (present + history) cost uniform in 0-15 added on a fraction of
the free cells. The via grid is blocked where any layer is, and costs 1 plus the mean congestion over the layers.
Searches: source / target pairs at uniform random free states, each solved alone over the whole grid.
"""
import numpy as np

from .abi import BLOCK, FREE, INF, Problem, default_dirw

ORTH, DIAG = 12, 17          # integer move weights: 12 per cell step, 17 ~ 12 * sqrt(2) per diagonal step
VIA = 50 * 12                # a via costs as much as 50 cells of track


def blob_mask(rng, ny, nx, frac, lmin=3, lmax=30):
    """Boolean (ny, nx) mask of random-walk blobs of lmin..lmax cells covering about frac of the cells."""
    m = np.zeros((ny, nx), bool)
    target = frac * nx * ny
    steps = np.array([(1, 0), (-1, 0), (0, 1), (0, -1)], np.int64)
    while m.sum() < target:
        need = target - m.sum()
        nb = max(16, int(need / ((lmin + lmax) / 2)) + 1)
        L = rng.integers(lmin, lmax + 1, nb)
        x = rng.integers(0, nx, nb)
        y = rng.integers(0, ny, nb)
        d = steps[rng.integers(0, 4, (nb, lmax - 1))]
        px = np.concatenate([x[:, None], x[:, None] + np.cumsum(d[:, :, 0], 1)], 1)
        py = np.concatenate([y[:, None], y[:, None] + np.cumsum(d[:, :, 1], 1)], 1)
        keep = np.arange(lmax)[None, :] < L[:, None]
        px = np.clip(px[keep], 0, nx - 1)
        py = np.clip(py[keep], 0, ny - 1)
        m[py, px] = True
    return m


def kernel_board(seed, nx=2000, ny=1600, nl=4, obstacles=0.20, congested=0.30, cmax=15):
    """(cst, st, vcst, vst) of a synthetic board."""
    rng = np.random.default_rng(seed)
    cst = np.ones((nl, ny, nx), np.int32)
    st = np.full((nl, ny, nx), FREE, np.uint16)
    extra_sum = np.zeros((ny, nx), np.int64)
    anyblk = np.zeros((ny, nx), bool)
    for l in range(nl):
        blk = blob_mask(rng, ny, nx, obstacles)
        extra = np.where(rng.random((ny, nx)) < congested, rng.integers(0, cmax + 1, (ny, nx)), 0)
        cst[l] = 1 + extra
        cst[l][blk] = INF
        st[l][blk] = BLOCK
        extra_sum += np.where(blk, 0, extra)
        anyblk |= blk
    vcst = (1 + extra_sum // nl).astype(np.int32)
    vst = np.full((ny, nx), FREE, np.uint16)
    vcst[anyblk] = INF
    vst[anyblk] = BLOCK
    return cst, st, vcst, vst


def free_states(st, rng, n):
    """n uniform random free states as (l, i, j)."""
    nl, ny, nx = st.shape
    out = []
    while len(out) < n:
        l, i, j = int(rng.integers(nl)), int(rng.integers(nx)), int(rng.integers(ny))
        if st[l, j, i] == FREE:
            out.append((l, i, j))
    return out


def pair_problems(board, seed, n, via_cost=VIA, orth=ORTH, diag=DIAG):
    """n single source / single target searches over the whole grid of board."""
    cst, st, vcst, vst = board
    nl, ny, nx = cst.shape
    rng = np.random.default_rng(seed)
    src = free_states(st, rng, n)
    dst = free_states(st, rng, n)
    probs = []
    for (l0, i0, j0), (l1, i1, j1) in zip(src, dst):
        s = (l0 * ny + j0) * nx + i0
        t = (l1 * ny + j1) * nx + i1
        probs.append(Problem(cst, st, vcst, vst, starts=[(s, 0)], targets=[t], via_cost=via_cost,
                             dirw=default_dirw(nl, orth, diag)))
    return probs


def class_grids(seed, nx=2400, ny=1800, nl=6, obstacles=0.05, clearances=(1, 2), congested=0.30, cmax=15):
    """One board seen by several net classes: the same copper and keepouts (random blobs), grown by each class's
    clearance in cells (its track half width plus clearance), so every class has its own static and cost maps.
    Returns [(cst, st, vcst, vst)] per class."""
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    base = [blob_mask(rng, ny, nx, obstacles) for _ in range(nl)]
    extra = [np.where(rng.random((ny, nx)) < congested, rng.integers(0, cmax + 1, (ny, nx)), 0) for _ in range(nl)]
    out = []
    for c in clearances:
        k = np.ones((2 * c + 1, 2 * c + 1), bool)
        cst = np.empty((nl, ny, nx), np.int32)
        st = np.zeros((nl, ny, nx), np.uint16)
        anyblk = np.zeros((ny, nx), bool)
        esum = np.zeros((ny, nx), np.int64)
        for l in range(nl):
            blk = ndimage.binary_dilation(base[l], structure=k)
            cst[l] = 1 + extra[l]
            cst[l][blk] = INF
            st[l][blk] = BLOCK
            anyblk |= blk
            esum += np.where(blk, 0, extra[l])
        vcst = (1 + esum // nl).astype(np.int32)
        vst = np.zeros((ny, nx), np.uint16)
        vcst[anyblk] = INF
        vst[anyblk] = BLOCK
        out.append((cst, st, vcst, vst))
    return out


def net_searches(grids, seed, n=256, smin=64, smax=2048, margin=40, via_cost=VIA, orth=ORTH, diag=DIAG):
    """n net searches of one speculative round: each net picks a class, a bounding box with sides log-uniform in
    [smin, smax] cells and 2-8 pins at free states inside it; its search runs from the first pin to the others
    inside the box grown by margin. Returns [(class index, Problem)]."""
    rng = np.random.default_rng(seed)
    nl, ny, nx = grids[0][0].shape
    out = []
    while len(out) < n:
        k = int(rng.integers(len(grids)))
        cst, st, vcst, vst = grids[k]
        bw = int(min(nx, np.exp(rng.uniform(np.log(smin), np.log(smax)))))
        bh = int(min(ny, np.exp(rng.uniform(np.log(smin), np.log(smax)))))
        x0, y0 = int(rng.integers(0, nx - bw + 1)), int(rng.integers(0, ny - bh + 1))
        npins = int(rng.integers(2, 9))
        pins = []
        for _ in range(50 * npins):
            if len(pins) == npins:
                break
            l, i, j = int(rng.integers(nl)), x0 + int(rng.integers(bw)), y0 + int(rng.integers(bh))
            if st[l, j, i] == FREE:
                pins.append((l, i, j))
        if len(pins) < 2:
            continue
        wi0, wj0 = max(0, x0 - margin), max(0, y0 - margin)
        wi1, wj1 = min(nx, x0 + bw + margin), min(ny, y0 + bh + margin)
        ww, wh = wi1 - wi0, wj1 - wj0

        def ws(p):
            l, i, j = p
            return (l * wh + (j - wj0)) * ww + (i - wi0)

        out.append((k, Problem(cst, st, vcst, vst, starts=[(ws(pins[0]), 0)], targets=[ws(p) for p in pins[1:]],
                               window=(wi0, wj0, ww, wh), via_cost=via_cost, dirw=default_dirw(nl, orth, diag))))
    return out
