"""Seeded search problems for the parity tests: small boards with every feature of a query switched on at random
(window offsets, own copper, own lists, the prior with its corridor, unusable layers, no vias, several starts and
targets)."""
import numpy as np

from pcbroute_amd import synth
from pcbroute_amd.abi import BLOCK, FREE, INF, MAX_LAYERS, Problem


def random_case(seed, nx=96, ny=80, nl=3, features=True, sym=False):
    rng = np.random.default_rng(seed)
    cst, st, vcst, vst = synth.kernel_board(seed, nx=nx, ny=ny, nl=nl)
    me = 7
    if features:
        # copper of this net (negative cost, st == me) and of another net (st == 9)
        for owner in (me, 9):
            for _ in range(30):
                l, i, j = int(rng.integers(nl)), int(rng.integers(nx)), int(rng.integers(ny))
                if st[l, j, i] == FREE:
                    st[l, j, i] = owner
                    cst[l, j, i] = -int(rng.integers(1, 20))
            for _ in range(6):
                i, j = int(rng.integers(nx)), int(rng.integers(ny))
                if vst[j, i] == FREE:
                    vst[j, i] = owner
                    vcst[j, i] = -int(rng.integers(1, 5))
    if features and rng.random() < 0.7:
        ww, wh = int(rng.integers(nx // 2, nx + 1)), int(rng.integers(ny // 2, ny + 1))
        wi0, wj0 = int(rng.integers(0, nx - ww + 1)), int(rng.integers(0, ny - wh + 1))
    else:
        wi0, wj0, ww, wh = 0, 0, nx, ny
    dirw = np.zeros((MAX_LAYERS, 8), np.int32)
    for l in range(nl):
        if sym:  # the router's shape: +x/-x alike, +y/-y alike, the four diagonals alike
            wx, wy, wd = (int(v) for v in rng.integers(5, 40, 3))
            dirw[l] = [wx, wx, wy, wy, wd + 5, wd + 5, wd + 5, wd + 5]
        else:
            dirw[l, :4] = rng.integers(5, 40, 4)
            dirw[l, 4:] = rng.integers(7, 60, 4)
    lay = np.ones(nl, np.uint8)
    if features and nl > 1 and rng.random() < 0.3:
        lay[int(rng.integers(nl))] = 0
    via_cost = int(rng.integers(1, 400)) if (not features or rng.random() < 0.85) else -1

    def ws(l, i, j):
        return (l * wh + j) * ww + i

    starts = []
    for _ in range(int(rng.integers(1, 4))):
        l, i, j = int(rng.integers(nl)), int(rng.integers(ww)), int(rng.integers(wh))
        starts.append((ws(l, i, j), int(rng.integers(0, 60))))
    targets = []
    for _ in range(int(rng.integers(1, 6))):
        l, i, j = int(rng.integers(nl)), int(rng.integers(ww)), int(rng.integers(wh))
        targets.append(ws(l, i, j))
    own_cf, own_vcf, prior = [], [], None
    if features:
        for _ in range(int(rng.integers(0, 20))):
            l, i, j = int(rng.integers(nl)), int(rng.integers(ww)), int(rng.integers(wh))
            own_cf.append((ws(l, i, j), int(rng.choice([1, 2, 3, 50, INF]))))
        for _ in range(int(rng.integers(0, 8))):
            i, j = int(rng.integers(ww)), int(rng.integers(wh))
            own_vcf.append((j * ww + i, int(rng.choice([1, 2, 9, INF]))))
        if rng.random() < 0.6:
            px = 4
            pw, ph = (nx + px - 1) // px, (ny + px - 1) // px
            pi0, pj0 = int(rng.integers(-2, 3)), int(rng.integers(-2, 3))
            play = [int(rng.integers(-1, nl)) for _ in range(nl)] + [-1] * (MAX_LAYERS - nl)
            prior = dict(data=rng.integers(0, 256, (nl, ph, pw)).astype(np.uint8), pi0=pi0, pj0=pj0,
                         pcol=np.arange(nx) // px, prow=np.arange(ny) // px, play=play,
                         mlbias=np.clip(8 - np.arange(256) // 32, 0, 8), mthr=int(rng.integers(0, 120)),
                         corridor=bool(rng.random() < 0.5))
    return Problem(cst, st, vcst, vst, starts=starts, targets=targets, window=(wi0, wj0, ww, wh), me=me,
                   via_cost=via_cost, dirw=dirw, lay=lay, own_cf=own_cf, own_vcf=own_vcf, prior=prior)
