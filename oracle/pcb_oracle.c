/* pcb_oracle.c - CPU restatement of pcbroute's grid-search back end on integer costs.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP back end. Only tests, the smoke check and the benchmark's
 * CPU-baseline leg load it; the product library never does.
 *
 * What each part follows (paths relative to the pcbroute sources):
 *   resolve_state  src/gpu.cu:78-105   (prep_cf: cost map, own copper, prior corridor and factor, via landing)
 *   resolve_via    src/gpu.cu:107-115  (prep_v)
 *   build_rect     src/gpu.cu:440-480  (solve_rect: own lists written over the maps, then resolved; starts)
 *   solve_rect     the fixpoint of src/gpu.cu:157-363 (scans, vias, bound), computed by Dijkstra
 *   sweep_rect     the pass order of src/gpu.cu:488-524 (8 directions in code order, vias, bound, until no change)
 *   orc_solve      src/gpu.cu:575-619  (rectangle widening by the octile lower bound)
 *   orc_backtrace  src/router.cpp:1602-1679 (first strict minimum, moves 0..7 then vias 0..NL-1)
 *   orc_astar      src/router.cpp:1396-1781 (the CPU A* the back end finishes; octile box heuristic)
 * Integer restatement: move cost dirw[l][dm] * cf(q), via cost via_cost * vf(cell), relaxation kept only while the
 * new cost is <= the best target cost so far, and the canonical field described in include/pcbgpu.h.
 */
#include "pcb_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static const int DX[8] = {1, -1, 0, 0, 1, 1, -1, -1};
static const int DY[8] = {0, 0, 1, -1, 1, -1, 1, -1};

#define INF PCBGPU_INF

static int tbit_at(const pcbgpu_query* q, size_t ws) { return (int)((q->tbit[ws >> 6] >> (ws & 63)) & 1u); }

static int32_t clampcf(int64_t r) { return r < 1 ? 1 : r > PCBGPU_CF_MAX ? PCBGPU_CF_MAX : (int32_t)r; }

static int query_ok(const pcbgpu_query* q) {
    if (!q || q->nl < 1 || q->nl > PCBGPU_MAX_LAYERS || q->ww < 1 || q->wh < 1) return 0;
    if (q->wi0 < 0 || q->wj0 < 0 || q->wi0 + q->ww > q->nx || q->wj0 + q->wh > q->ny) return 0;
    if (!q->tbit || !q->vcst || !q->vst) return 0;
    if (q->via_cost == 0 || q->via_cost > PCBGPU_W_MAX) return 0;
    for (int l = 0; l < q->nl; ++l) {
        if (!q->cst[l] || !q->st[l]) return 0;
        if (!q->lay[l]) continue;
        for (int d = 0; d < 8; ++d)
            if (q->dirw[l][d] < 1 || q->dirw[l][d] > PCBGPU_W_MAX) return 0;
    }
    if (q->prior) {
        if (!q->pcol || !q->prow || q->pw < 0 || q->ph < 0) return 0;
        for (int l = 0; l < q->nl; ++l)
            if (q->lay[l] && q->play[l] >= q->pnl) return 0;
    }
    return 1;
}

/* prep_cf (src/gpu.cu:78-105) for one state: the cost factor (0: may not be entered) and whether a via may land */
static int32_t resolve_state(const pcbgpu_query* q, int l, int gi, int gj, int32_t v, uint16_t s, int tg, uint8_t* vok) {
    const int sok = s == PCBGPU_FREE || s == q->me;
    const int corr = q->prior && q->corridor;
    int64_t r = 0;
    int blocked = 0;
    if (!q->lay[l] || v >= INF) blocked = 1;
    else if (v < 0) { if (s == q->me) r = -(int64_t)v; else blocked = 1; }
    else r = v;
    int pv = 255;
    if (q->prior && q->lay[l]) {
        int pl = q->play[l];
        if (pl >= 0) {
            int x = q->pcol[gi] - q->pi0, y = q->prow[gj] - q->pj0;
            pv = (x < 0 || y < 0 || x >= q->pw || y >= q->ph) ? 0 : q->prior[((size_t)pl * q->ph + y) * q->pw + x];
        }
        if (corr && pv < q->mthr && !tg) blocked = 1;
        else if (!blocked) r += q->mlbias[pv];
    }
    *vok = (uint8_t)(q->lay[l] ? (sok && (!corr || tg || pv >= q->mthr)) : tg);
    return blocked ? 0 : clampcf(r);
}

/* prep_v (src/gpu.cu:107-115) for one cell */
static int32_t resolve_via(const pcbgpu_query* q, int32_t v, uint16_t s) {
    if (q->via_cost < 0 || v >= INF) return INF;
    int64_t r;
    if (v < 0) { if (s != q->me) return INF; r = -(int64_t)v; }
    else r = v;
    return (int32_t)(clampcf(r) * (int64_t)q->via_cost);
}

int orc_resolve(const pcbgpu_query* q, int rx0, int ry0, int rw, int rh, int32_t* cf, uint8_t* vok, uint8_t* tg,
                int32_t* vcf) {
    if (!query_ok(q) || rx0 < 0 || ry0 < 0 || rw < 1 || rh < 1 || rx0 + rw > q->ww || ry0 + rh > q->wh) return -1;
    const size_t RN = (size_t)rw * rh, N = RN * q->nl, WN = (size_t)q->ww * q->wh;
    /* the net's own cells: their value replaces the map's before it is resolved (src/gpu.cu:475-477); the last
     * entry of a state counts */
    int32_t* ov = (int32_t*)malloc(N * sizeof(int32_t));
    uint8_t* has = (uint8_t*)calloc(N, 1);
    int32_t* ovv = (int32_t*)malloc(RN * sizeof(int32_t));
    uint8_t* hasv = (uint8_t*)calloc(RN, 1);
    if (!ov || !has || !ovv || !hasv) { free(ov); free(has); free(ovv); free(hasv); return -1; }
    for (size_t k = 0; k < q->n_own_cf; ++k) {
        size_t ws = q->own_cf[k].state, l = ws / WN, c = ws % WN;
        if (l >= (size_t)q->nl) continue;
        int i = (int)(c % q->ww) - rx0, j = (int)(c / q->ww) - ry0;
        if (i < 0 || j < 0 || i >= rw || j >= rh) continue;
        size_t s = l * RN + (size_t)j * rw + i;
        ov[s] = q->own_cf[k].value; has[s] = 1;
    }
    for (size_t k = 0; k < q->n_own_vcf; ++k) {
        size_t c = q->own_vcf[k].state;
        if (c >= WN) continue;
        int i = (int)(c % q->ww) - rx0, j = (int)(c / q->ww) - ry0;
        if (i < 0 || j < 0 || i >= rw || j >= rh) continue;
        size_t s = (size_t)j * rw + i;
        ovv[s] = q->own_vcf[k].value; hasv[s] = 1;
    }
    for (int l = 0; l < q->nl; ++l)
        for (int j = 0; j < rh; ++j)
            for (int i = 0; i < rw; ++i) {
                const int gi = q->wi0 + rx0 + i, gj = q->wj0 + ry0 + j;
                const size_t cell = (size_t)gj * q->nx + gi, s = (size_t)l * RN + (size_t)j * rw + i;
                const size_t ws = ((size_t)l * q->wh + ry0 + j) * q->ww + rx0 + i;
                const int t = tbit_at(q, ws);
                int32_t v = has[s] ? ov[s] : q->cst[l][cell];
                cf[s] = resolve_state(q, l, gi, gj, v, q->st[l][cell], t, &vok[s]);
                tg[s] = (uint8_t)t;
            }
    for (int j = 0; j < rh; ++j)
        for (int i = 0; i < rw; ++i) {
            const size_t cell = (size_t)(q->wj0 + ry0 + j) * q->nx + q->wi0 + rx0 + i, s = (size_t)j * rw + i;
            vcf[s] = resolve_via(q, hasv[s] ? ovv[s] : q->vcst[cell], q->vst[cell]);
        }
    free(ov); free(has); free(ovv); free(hasv);
    return 0;
}

/* ---- binary min-heap of (cost << 32 | state) */
typedef struct { uint64_t* a; size_t n, cap; } Heap;
static int hpush(Heap* h, uint64_t x) {
    if (h->n == h->cap) {
        size_t nc = h->cap ? h->cap * 2 : 1024;
        uint64_t* na = (uint64_t*)realloc(h->a, nc * sizeof(uint64_t));
        if (!na) return -1;
        h->a = na; h->cap = nc;
    }
    size_t k = h->n++;
    while (k) {
        size_t p = (k - 1) / 2;
        if (h->a[p] <= x) break;
        h->a[k] = h->a[p]; k = p;
    }
    h->a[k] = x;
    return 0;
}
static uint64_t hpop(Heap* h) {
    uint64_t top = h->a[0], x = h->a[--h->n];
    size_t k = 0;
    for (;;) {
        size_t c = 2 * k + 1;
        if (c >= h->n) break;
        if (c + 1 < h->n && h->a[c + 1] < h->a[c]) ++c;
        if (x <= h->a[c]) break;
        h->a[k] = h->a[c]; k = c;
    }
    if (h->n) h->a[k] = x;
    return top;
}

typedef struct {
    int rw, rh, nl;
    size_t RN, N;
    int32_t *cf, *vcf;
    uint8_t *vok, *tg;
} Rect;

static void rect_free(Rect* R) { free(R->cf); free(R->vcf); free(R->vok); free(R->tg); }

static int rect_build(const pcbgpu_query* q, int rx0, int ry0, int rw, int rh, Rect* R) {
    memset(R, 0, sizeof *R);
    R->rw = rw; R->rh = rh; R->nl = q->nl;
    R->RN = (size_t)rw * rh; R->N = R->RN * q->nl;
    R->cf = (int32_t*)malloc(R->N * 4); R->vok = (uint8_t*)malloc(R->N); R->tg = (uint8_t*)malloc(R->N);
    R->vcf = (int32_t*)malloc(R->RN * 4);
    if (!R->cf || !R->vok || !R->tg || !R->vcf) { rect_free(R); return -1; }
    if (orc_resolve(q, rx0, ry0, rw, rh, R->cf, R->vok, R->tg, R->vcf)) { rect_free(R); return -1; }
    return 0;
}

/* starts (src/gpu.cu:478, src/router.cpp:1573-1584): the least g0 of a state; g0 >= INF is no start */
static void seed_starts(const pcbgpu_query* q, int rx0, int ry0, const Rect* R, int32_t* d) {
    const size_t WN = (size_t)q->ww * q->wh;
    for (size_t s = 0; s < R->N; ++s) d[s] = INF;
    for (size_t k = 0; k < q->n_starts; ++k) {
        size_t ws = q->starts[k].state, l = ws / WN, c = ws % WN;
        int32_t g = q->starts[k].value;
        if (l >= (size_t)q->nl || g >= INF) continue;
        if (g < 0) g = 0;
        int i = (int)(c % q->ww) - rx0, j = (int)(c / q->ww) - ry0;
        if (i < 0 || j < 0 || i >= R->rw || j >= R->rh) continue;
        size_t s = l * R->RN + (size_t)j * R->rw + i;
        if (g < d[s]) d[s] = g;
    }
}

/* opt, best (the lowest state among the cheapest targets) and the canonical field */
static void canonical(const Rect* R, int32_t* d, int64_t* best, int32_t* opt) {
    int32_t o = INF;
    int64_t b = -1;
    for (size_t s = 0; s < R->N; ++s)
        if (R->tg[s] && d[s] < o) { o = d[s]; b = (int64_t)s; }
    for (size_t s = 0; s < R->N; ++s)
        if (d[s] >= o) d[s] = ((int64_t)s == b) ? o : INF;
    *best = b; *opt = o;
}

int orc_solve_rect(const pcbgpu_query* q, int rx0, int ry0, int rw, int rh, int32_t* d, int64_t* best, int32_t* opt) {
    Rect R;
    if (rect_build(q, rx0, ry0, rw, rh, &R)) return -1;
    seed_starts(q, rx0, ry0, &R, d);
    Heap h = {0, 0, 0};
    for (size_t s = 0; s < R.N; ++s)
        if (d[s] < INF && hpush(&h, ((uint64_t)d[s] << 32) | s)) { free(h.a); rect_free(&R); return -1; }
    while (h.n) {
        uint64_t top = hpop(&h);
        const int32_t k = (int32_t)(top >> 32);
        const size_t s = (size_t)(top & 0xFFFFFFFFu);
        if (k != d[s]) continue;
        const int l = (int)(s / R.RN);
        const size_t c = s % R.RN;
        const int j = (int)(c / rw), i = (int)(c % rw);
        /* track moves out of s: into a neighbour on the same layer that may be entered */
        for (int dm = 0; dm < 8; ++dm) {
            const int ni = i + DX[dm], nj = j + DY[dm];
            if (ni < 0 || nj < 0 || ni >= rw || nj >= rh) continue;
            const size_t t = (size_t)l * R.RN + (size_t)nj * rw + ni;
            if (!R.cf[t]) continue;
            int64_t nd = (int64_t)k + (int64_t)q->dirw[l][dm] * R.cf[t];
            if (nd < d[t]) { d[t] = (int32_t)nd; if (hpush(&h, ((uint64_t)nd << 32) | t)) goto oom; }
        }
        /* vias at the cell to every other layer a via may land on */
        if (R.vcf[c] < INF)
            for (int l2 = 0; l2 < R.nl; ++l2) {
                if (l2 == l) continue;
                const size_t t = (size_t)l2 * R.RN + c;
                if (!R.vok[t]) continue;
                int64_t nd = (int64_t)k + R.vcf[c];
                if (nd < d[t]) { d[t] = (int32_t)nd; if (hpush(&h, ((uint64_t)nd << 32) | t)) goto oom; }
            }
    }
    free(h.a);
    canonical(&R, d, best, opt);
    rect_free(&R);
    return 0;
oom:
    free(h.a);
    rect_free(&R);
    return -1;
}

/* One directional scan (src/gpu.cu:157-329 restated as plain sequential line walks): every line of direction dm on
 * every layer, from its first cell on. Returns 1 if anything changed. */
static int scan_dir(const pcbgpu_query* q, const Rect* R, int32_t* d, int dm, int64_t ub) {
    const int rw = R->rw, rh = R->rh, dx = DX[dm], dy = DY[dm];
    int changed = 0;
    for (int l = 0; l < R->nl; ++l) {
        if (!q->lay[l]) continue;
        int32_t* D = d + (size_t)l * R->RN;
        const int32_t* C = R->cf + (size_t)l * R->RN;
        const int64_t w = q->dirw[l][dm];
        /* line starts: the entry column (dx != 0) and the entry row (dy != 0); the shared corner once */
        const int nlines = (dx ? rh : 0) + (dy ? rw : 0);
        for (int t = 0; t < nlines; ++t) {
            int i, j;
            if (dx && t < rh) { i = dx > 0 ? 0 : rw - 1; j = t; }
            else {
                i = dx ? t - rh : t;
                j = dy > 0 ? 0 : rh - 1;
                if (dx && i == (dx > 0 ? 0 : rw - 1)) continue;
            }
            int64_t prev = INF;
            for (; i >= 0 && j >= 0 && i < rw && j < rh; i += dx, j += dy) {
                const size_t s = (size_t)j * rw + i;
                int32_t cur = D[s];
                if (prev < INF && C[s]) {
                    int64_t cand = prev + w * C[s];
                    if (cand < cur && cand <= ub) { cur = (int32_t)cand; D[s] = cur; changed = 1; }
                }
                prev = cur;
            }
        }
    }
    return changed;
}

int orc_sweep_rect(const pcbgpu_query* q, int rx0, int ry0, int rw, int rh, int32_t* d, int64_t* best, int32_t* opt) {
    Rect R;
    if (rect_build(q, rx0, ry0, rw, rh, &R)) return -1;
    seed_starts(q, rx0, ry0, &R, d);
    int64_t ub = INF - 1;
    int sweeps = 0;
    for (;; ++sweeps) {
        int changed = 0;
        for (int dm = 0; dm < 8; ++dm) changed |= scan_dir(q, &R, d, dm, ub);
        if (q->via_cost >= 0)
            for (size_t c = 0; c < R.RN; ++c) {
                if (R.vcf[c] >= INF) continue;
                int32_t b = INF;
                for (int l = 0; l < R.nl; ++l) if (d[l * R.RN + c] < b) b = d[l * R.RN + c];
                if (b >= INF) continue;
                for (int l = 0; l < R.nl; ++l) {
                    const size_t s = l * R.RN + c;
                    int64_t cand = (int64_t)b + R.vcf[c];
                    if (R.vok[s] && cand < d[s] && cand <= ub) { d[s] = (int32_t)cand; changed = 1; }
                }
            }
        for (size_t s = 0; s < R.N; ++s)
            if (R.tg[s] && d[s] < ub) ub = d[s];
        if (!changed) break;
    }
    canonical(&R, d, best, opt);
    rect_free(&R);
    return sweeps + 1;
}

/* ---- rectangle widening (src/gpu.cu:575-619) */
typedef struct { int x0, y0, x1, y1; } Box;

static void box_add(Box* b, int i, int j) {
    if (i < b->x0) b->x0 = i;
    if (i > b->x1) b->x1 = i;
    if (j < b->y0) b->y0 = j;
    if (j > b->y1) b->y1 = j;
}

static int64_t oct_to(int i, int j, const Box* b, int64_t cpo, int64_t cpd) {
    int64_t dx = b->x0 - i > 0 ? b->x0 - i : (i - b->x1 > 0 ? i - b->x1 : 0);
    int64_t dy = b->y0 - j > 0 ? b->y0 - j : (j - b->y1 > 0 ? j - b->y1 : 0);
    int64_t m = dx < dy ? dx : dy, M = dx < dy ? dy : dx;
    return m * cpd + (M - m) * cpo;
}

/* the least orthogonal and diagonal step costs over the usable layers: every path of a track from a box to a cell
 * at octile offset (dx, dy) costs at least min(dx,dy) * cpd + |dx - dy| * cpo (cost factors are >= 1) */
static void step_floor(const pcbgpu_query* q, int64_t* cpo, int64_t* cpd) {
    int64_t wo = PCBGPU_W_MAX, wd = PCBGPU_W_MAX;
    for (int l = 0; l < q->nl; ++l) {
        if (!q->lay[l]) continue;
        for (int d = 0; d < 4; ++d) if (q->dirw[l][d] < wo) wo = q->dirw[l][d];
        for (int d = 4; d < 8; ++d) if (q->dirw[l][d] < wd) wd = q->dirw[l][d];
    }
    *cpo = wo < wd ? wo : wd;
    *cpd = wd < 2 * wo ? wd : 2 * wo;
}

int orc_solve(const pcbgpu_query* q, pcbgpu_out* out, size_t max_states) {
    memset(out, 0, sizeof *out);
    out->best = -1;
    if (!query_ok(q)) return -1;
    const size_t WN = (size_t)q->ww * q->wh, WS = WN * q->nl, WB = WS / 64 + 1;
    Box sb = {1 << 30, 1 << 30, -1, -1}, tb = {1 << 30, 1 << 30, -1, -1};
    for (size_t k = 0; k < q->n_starts; ++k) {
        if (q->starts[k].state >= WS) continue;
        size_t c = q->starts[k].state % WN;
        int i = (int)(c % q->ww), j = (int)(c / q->ww);
        box_add(&sb, i, j);
    }
    for (size_t w = 0; w < WB; ++w) {
        uint64_t b = q->tbit[w];
        while (b) {
            int k = __builtin_ctzll(b);
            b &= b - 1;
            size_t st = w * 64 + k;
            if (st >= WS) continue;
            size_t c = st % WN;
            int i = (int)(c % q->ww), j = (int)(c / q->ww);
            box_add(&tb, i, j);
        }
    }
    if (sb.x1 < 0 || tb.x1 < 0) return -1;
    int64_t cpo, cpd;
    step_floor(q, &cpo, &cpd);
    const int hx0 = sb.x0 < tb.x0 ? sb.x0 : tb.x0, hy0 = sb.y0 < tb.y0 ? sb.y0 : tb.y0;
    const int hx1 = sb.x1 > tb.x1 ? sb.x1 : tb.x1, hy1 = sb.y1 > tb.y1 ? sb.y1 : tb.y1;
    int32_t* d = NULL;
    int64_t M = 80;
    for (int tries = 0;; M *= 3, ++tries) {
        int64_t rx0 = hx0 - M, ry0 = hy0 - M, rx1 = hx1 + M + 1, ry1 = hy1 + M + 1;
        if (rx0 < 0) rx0 = 0;
        if (ry0 < 0) ry0 = 0;
        if (rx1 > q->ww) rx1 = q->ww;
        if (ry1 > q->wh) ry1 = q->wh;
        if (tries == 0 && 2 * (rx1 - rx0) * (ry1 - ry0) > (int64_t)q->ww * q->wh) { rx0 = ry0 = 0; rx1 = q->ww; ry1 = q->wh; }
        const int full = rx0 == 0 && ry0 == 0 && rx1 == q->ww && ry1 == q->wh;
        const int rw = (int)(rx1 - rx0), rh = (int)(ry1 - ry0);
        const size_t N = (size_t)rw * rh * q->nl;
        if (max_states && N > max_states) { free(d); return -1; }
        free(d);
        d = (int32_t*)malloc(N * 4);
        if (!d) return -1;
        int64_t best;
        int32_t opt;
        if (orc_solve_rect(q, (int)rx0, (int)ry0, rw, rh, d, &best, &opt)) { free(d); return -1; }
        out->rx0 = (int)rx0; out->ry0 = (int)ry0; out->rw = rw; out->rh = rh;
        out->nl = q->nl; out->ww = q->ww; out->wh = q->wh;
        out->d = d; out->cap = N; out->tries = tries + 1;
        out->best = -1;
        if (best >= 0) {
            const size_t RN = (size_t)rw * rh, l = (size_t)best / RN, c = (size_t)best % RN;
            out->best = (int64_t)((l * q->wh + c / rw + ry0) * q->ww + c % rw + rx0);
        }
        if (full) break;
        if (opt < INF) {
            int64_t mn = INT64_MAX;
            if (rx0 > 0) for (int j = 0; j < q->wh; ++j) { int64_t v = oct_to((int)rx0 - 1, j, &sb, cpo, cpd) + oct_to((int)rx0 - 1, j, &tb, cpo, cpd); if (v < mn) mn = v; }
            if (rx1 < q->ww) for (int j = 0; j < q->wh; ++j) { int64_t v = oct_to((int)rx1, j, &sb, cpo, cpd) + oct_to((int)rx1, j, &tb, cpo, cpd); if (v < mn) mn = v; }
            if (ry0 > 0) for (int i = 0; i < q->ww; ++i) { int64_t v = oct_to(i, (int)ry0 - 1, &sb, cpo, cpd) + oct_to(i, (int)ry0 - 1, &tb, cpo, cpd); if (v < mn) mn = v; }
            if (ry1 < q->wh) for (int i = 0; i < q->ww; ++i) { int64_t v = oct_to(i, (int)ry1, &sb, cpo, cpd) + oct_to(i, (int)ry1, &tb, cpo, cpd); if (v < mn) mn = v; }
            if (mn > opt) break;
        }
    }
    return 0;
}

void orc_out_free(pcbgpu_out* out) {
    free(out->d);
    out->d = NULL;
    out->cap = 0;
}

/* ---- backtrace (src/router.cpp:1602-1679) */
static int32_t field_at(const pcbgpu_out* o, int64_t ws) {
    const int64_t WN = (int64_t)o->ww * o->wh;
    const int l = (int)(ws / WN);
    const int64_t c = ws % WN;
    const int i = (int)(c % o->ww) - o->rx0, j = (int)(c / o->ww) - o->ry0;
    if (i < 0 || j < 0 || i >= o->rw || j >= o->rh) return INF;
    return o->d[((size_t)l * o->rh + j) * o->rw + i];
}

int64_t orc_backtrace(const pcbgpu_query* q, const pcbgpu_out* out, int32_t via_gapc, double via_gap2,
                      const uint32_t* old_vias, size_t n_old_vias, uint32_t* chain_states, uint8_t* chain_codes,
                      size_t chain_cap) {
    if (!query_ok(q) || !out || !out->d || out->best < 0) return PCBGPU_ERR_FAIL;
    Rect R;
    if (rect_build(q, out->rx0, out->ry0, out->rw, out->rh, &R)) return PCBGPU_ERR_FAIL;
    const int64_t WN = (int64_t)q->ww * q->wh, NS = WN * q->nl;
    /* the start costs: the least g0 of each start state */
    int32_t* g0 = (int32_t*)malloc(R.N * 4);
    if (!g0) { rect_free(&R); return PCBGPU_ERR_FAIL; }
    seed_starts(q, out->rx0, out->ry0, &R, g0);
    int64_t n = 0, rc = 0;
    int64_t since_via = 1 << 20;
    int64_t qs = out->best;
    uint32_t* vias = NULL;
    size_t nv = 0;
    for (int64_t guard = 0; guard < NS; ++guard) {
        const int l = (int)(qs / WN);
        const int64_t c = qs % WN;
        const int i = (int)(c % q->ww), j = (int)(c / q->ww);
        const int ri = i - out->rx0, rj = j - out->ry0;
        if (ri < 0 || rj < 0 || ri >= out->rw || rj >= out->rh) { rc = PCBGPU_ERR_CHAIN; goto done; }
        const size_t rs = ((size_t)l * out->rh + rj) * out->rw + ri;
        const int32_t dq = field_at(out, qs);
        if ((size_t)n >= chain_cap) { rc = PCBGPU_ERR_FAIL; goto done; }
        if (g0[rs] < INF && dq >= g0[rs]) { chain_states[n] = (uint32_t)qs; chain_codes[n] = 255; ++n; break; }
        int64_t bc = INT64_MAX, bp = -1;
        int code = 0;
        if (q->lay[l] && R.cf[rs]) {
            for (int dm = 0; dm < 8; ++dm) {
                const int pi = i - DX[dm], pj = j - DY[dm];
                if (pi < 0 || pj < 0 || pi >= q->ww || pj >= q->wh) continue;
                const int64_t p = l * WN + (int64_t)pj * q->ww + pi;
                const int64_t cst = (int64_t)field_at(out, p) + (int64_t)q->dirw[l][dm] * R.cf[rs];
                if (cst < bc) { bc = cst; bp = p; code = dm; }
            }
        }
        if (q->via_cost >= 0 && R.vok[rs] && R.vcf[(size_t)rj * out->rw + ri] < INF) {
            const int64_t vc = R.vcf[(size_t)rj * out->rw + ri];
            for (int l2 = 0; l2 < q->nl; ++l2) {
                if (l2 == l) continue;
                const int64_t p = l2 * WN + c;
                const int64_t cst = (int64_t)field_at(out, p) + vc;
                if (cst < bc) { bc = cst; bp = p; code = 8 + l2; }
            }
        }
        if (bp < 0 || bc > dq) { rc = PCBGPU_ERR_CHAIN; goto done; }
        chain_states[n] = (uint32_t)qs; chain_codes[n] = (uint8_t)code; ++n;
        if (code >= 8) {
            if (since_via < via_gapc) { rc = PCBGPU_ERR_CHAIN; goto done; }
            since_via = 0;
            uint32_t* nvp = (uint32_t*)realloc(vias, (nv + 1) * 4);
            if (!nvp) { rc = PCBGPU_ERR_FAIL; goto done; }
            vias = nvp;
            vias[nv++] = (uint32_t)((size_t)(j + q->wj0) * q->nx + (i + q->wi0));
        } else ++since_via;
        qs = bp;
    }
    if (n == 0 || chain_codes[n - 1] != 255) { rc = PCBGPU_ERR_CHAIN; goto done; }
    for (size_t a = 0; a < nv; ++a)
        for (size_t b = 0; b < n_old_vias; ++b) {
            const double ddx = (double)(old_vias[b] % q->nx) - (double)(vias[a] % q->nx);
            const double ddy = (double)(old_vias[b] / q->nx) - (double)(vias[a] / q->nx);
            if (old_vias[b] != vias[a] && ddx * ddx + ddy * ddy < via_gap2) { rc = PCBGPU_ERR_CHAIN; goto done; }
        }
    rc = n;
done:
    free(vias);
    free(g0);
    rect_free(&R);
    return rc;
}

/* ---- A* from the starts to the targets over the whole window (the CPU search; src/router.cpp:1682-1770) */
int32_t orc_astar(const pcbgpu_query* q, int64_t* pops) {
    *pops = 0;
    if (!query_ok(q)) return INF;
    Rect R;
    if (rect_build(q, 0, 0, q->ww, q->wh, &R)) return INF;
    const int rw = q->ww, rh = q->wh;
    Box tb = {1 << 30, 1 << 30, -1, -1};
    for (size_t s = 0; s < R.N; ++s)
        if (R.tg[s]) {
            const int i = (int)((s % R.RN) % rw), j = (int)((s % R.RN) / rw);
            box_add(&tb, i, j);
        }
    int32_t* g = (int32_t*)malloc(R.N * 4);
    uint8_t* closed = (uint8_t*)calloc(R.N, 1);
    if (!g || !closed || tb.x1 < 0) { free(g); free(closed); rect_free(&R); return INF; }
    int64_t cpo, cpd;
    step_floor(q, &cpo, &cpd);
    seed_starts(q, 0, 0, &R, g);
    Heap h = {0, 0, 0};
    for (size_t s = 0; s < R.N; ++s)
        if (g[s] < INF) {
            const int i = (int)((s % R.RN) % rw), j = (int)((s % R.RN) / rw);
            hpush(&h, ((uint64_t)(g[s] + oct_to(i, j, &tb, cpo, cpd)) << 32) | s);
        }
    int32_t res = INF;
    while (h.n) {
        uint64_t top = hpop(&h);
        const size_t s = (size_t)(top & 0xFFFFFFFFu);
        if (closed[s]) continue;
        closed[s] = 1;
        ++*pops;
        if (R.tg[s]) { res = g[s]; break; }
        const int l = (int)(s / R.RN);
        const size_t c = s % R.RN;
        const int j = (int)(c / rw), i = (int)(c % rw);
        for (int dm = 0; dm < 8; ++dm) {
            const int ni = i + DX[dm], nj = j + DY[dm];
            if (ni < 0 || nj < 0 || ni >= rw || nj >= rh) continue;
            const size_t t = (size_t)l * R.RN + (size_t)nj * rw + ni;
            if (!R.cf[t] || closed[t]) continue;
            const int64_t nd = (int64_t)g[s] + (int64_t)q->dirw[l][dm] * R.cf[t];
            if (nd < g[t]) { g[t] = (int32_t)nd; hpush(&h, ((uint64_t)(nd + oct_to(ni, nj, &tb, cpo, cpd)) << 32) | t); }
        }
        if (R.vcf[c] < INF)
            for (int l2 = 0; l2 < R.nl; ++l2) {
                const size_t t = (size_t)l2 * R.RN + c;
                if (l2 == l || !R.vok[t] || closed[t]) continue;
                const int64_t nd = (int64_t)g[s] + R.vcf[c];
                if (nd < g[t]) { g[t] = (int32_t)nd; hpush(&h, ((uint64_t)(nd + oct_to(i, j, &tb, cpo, cpd)) << 32) | t); }
            }
    }
    free(h.a); free(g); free(closed);
    rect_free(&R);
    return res;
}

typedef struct {
    const pcbgpu_query* qs;
    int n, nthreads, id;
    int32_t* opt;
    int64_t* pops;
} Task;

static void* astar_worker(void* a) {
    Task* t = (Task*)a;
    for (int k = t->id; k < t->n; k += t->nthreads) t->opt[k] = orc_astar(&t->qs[k], &t->pops[k]);
    return NULL;
}

int orc_astar_batch(const pcbgpu_query* qs, int n, int nthreads, int32_t* opt, int64_t* pops) {
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    Task* ts = (Task*)malloc(sizeof(Task) * nthreads);
    for (int k = 0; k < nthreads; ++k) {
        ts[k] = (Task){qs, n, nthreads, k, opt, pops};
        pthread_create(&th[k], NULL, astar_worker, &ts[k]);
    }
    for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    free(th); free(ts);
    return 0;
}
