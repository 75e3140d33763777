// astar.cpp - the CPU search (astar.h): A* over a query's window with a canonical finish.
//
// Moves and costs are those of the device search (src/router.cpp:1725-1769 on integer costs): a track move dm from
// p into q on layer l costs dirw[l][dm] * cf(q) where q may be entered, a via from any layer to layer l2 at a cell
// costs the cell's via cost where a via may land on l2. States are window states l * ww * wh + j * ww + i.
#include "astar.h"

#include "backtrace.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <queue>

namespace pcb {

namespace {

constexpr int DXS[8] = {1, -1, 0, 0, 1, 1, -1, -1};
constexpr int DYS[8] = {0, 0, 1, -1, 1, -1, 1, -1};

// the heap orders by f, then deeper states (larger g) first
struct ItemLess {
    template <class T>
    bool operator()(const T& a, const T& b) const { return a.f != b.f ? a.f > b.f : a.g < b.g; }
};

}  // namespace

// The cheapest walk over an empty grid covering dx columns and dy rows: min(dx, dy) diagonal steps and the rest
// straight, or all straight where two straight steps cost less than a diagonal one.
static inline int64_t walk(int64_t dx, int64_t dy, int64_t wx, int64_t wy, int64_t wd) {
    const int64_t m = std::min(dx, dy);
    return m * wd + (dx - m) * wx + (dy - m) * wy;
}

int32_t Astar::h(int l, int i, int j) const {
    const int64_t wd = std::min<int64_t>(wd_, int64_t(wx_) + wy_);
    if (hk_ == COARSE) {
        const int32_t d = coarse_[size_t(j / CS) * cw_ + (i / CS)];
        if (d >= INF) return INF;
        return int32_t(std::max<int64_t>(0, int64_t(CS) * d - int64_t(wmax_) * (CS - 1)));
    }
    int64_t best = INF;
    for (const TargetBox& b : boxes_) {
        const bool other = b.l >= 0 && b.l != l;
        if (other && vmin_ >= INF) continue;
        const int64_t dx = std::max({int64_t(b.i0) - i, int64_t(i) - b.i1, int64_t(0)});
        const int64_t dy = std::max({int64_t(b.j0) - j, int64_t(j) - b.j1, int64_t(0)});
        best = std::min(best, walk(dx, dy, wx_, wy_, wd) + (other ? vmin_ : 0));
    }
    return int32_t(std::min<int64_t>(best, INF));
}

// The distance map over blocks of CS x CS cells: blocks holding a target are 0, a step to a neighbouring block costs
// the least move weight of its direction (an empty grid), Dijkstra over the blocks.
void Astar::build_coarse() {
    const pcbgpu_query* q = qc_.q;
    cw_ = (q->ww + CS - 1) / CS;
    ch_ = (q->wh + CS - 1) / CS;
    coarse_.assign(size_t(cw_) * ch_, INF);
    struct Node {
        int32_t d;
        uint32_t c;
        bool operator>(const Node& o) const { return d > o.d; }
    };
    std::priority_queue<Node, std::vector<Node>, std::greater<Node>> pq;
    const int64_t words = (qc_.NS + 63) / 64;
    for (int64_t w = 0; w < words; ++w) {
        uint64_t bits = q->tbit[w];
        while (bits) {
            const int64_t ws = w * 64 + __builtin_ctzll(bits);
            bits &= bits - 1;
            if (ws >= qc_.NS) break;
            const int64_t c = ws % qc_.WN;
            const uint32_t cb = uint32_t((c / q->ww) / CS * cw_ + (c % q->ww) / CS);
            if (coarse_[cb] != 0) { coarse_[cb] = 0; pq.push({0, cb}); }
        }
    }
    const int32_t wd = int32_t(std::min<int64_t>(wd_, int64_t(wx_) + wy_));
    const int32_t wstep[8] = {wx_, wx_, wy_, wy_, wd, wd, wd, wd};
    while (!pq.empty()) {
        const Node n = pq.top();
        pq.pop();
        if (n.d > coarse_[n.c]) continue;
        const int ci = int(n.c % cw_), cj = int(n.c / cw_);
        for (int dm = 0; dm < 8; ++dm) {
            const int ni = ci + DXS[dm], nj = cj + DYS[dm];
            if (ni < 0 || nj < 0 || ni >= cw_ || nj >= ch_) continue;
            const uint32_t nc = uint32_t(nj * cw_ + ni);
            const int32_t nd = n.d + wstep[dm];
            if (nd < coarse_[nc]) { coarse_[nc] = nd; pq.push({nd, nc}); }
        }
    }
}

bool Astar::begin(const pcbgpu_query* q, Heuristic hk, const TargetBox* boxes, size_t nbox, bool reverse,
                  int32_t cf_min) {
    best = -1;
    opt = INF;
    pops = 0;
    found_ = false;
    broken_ = false;
    cfmin_ = std::clamp<int32_t>(cf_min, 1, PCBGPU_CF_MAX);
    rules_ = false;
    heap_.clear();
    if (!qc_.init(q)) return false;
    if (q->wi0 < 0 || q->wj0 < 0 || int64_t(q->wi0) + q->ww > q->nx || int64_t(q->wj0) + q->wh > q->ny) return false;
    if (q->via_cost == 0 || q->via_cost > PCBGPU_W_MAX) return false;
    hk_ = hk;
    reverse_ = reverse;
    // the least weight of each move direction over the layers a track may use (none: no track moves, h = 0)
    wx_ = wy_ = wd_ = PCBGPU_W_MAX;
    bool any = false;
    for (int l = 0; l < q->nl; ++l) {
        if (!q->lay[l]) continue;
        for (int d = 0; d < 8; ++d)
            if (q->dirw[l][d] < 1 || q->dirw[l][d] > PCBGPU_W_MAX) return false;
        any = true;
        wx_ = std::min({wx_, q->dirw[l][0], q->dirw[l][1]});
        wy_ = std::min({wy_, q->dirw[l][2], q->dirw[l][3]});
        wd_ = std::min({wd_, q->dirw[l][4], q->dirw[l][5], q->dirw[l][6], q->dirw[l][7]});
    }
    if (!any) wx_ = wy_ = wd_ = 0;
    // every step costs at least its weight times the least factor (weights <= 1023, factors < 2^20: no overflow in
    // the 64-bit sums of h; the products are capped below INF)
    wx_ = int32_t(std::min<int64_t>(int64_t(wx_) * cfmin_, INF - 1));
    wy_ = int32_t(std::min<int64_t>(int64_t(wy_) * cfmin_, INF - 1));
    wd_ = int32_t(std::min<int64_t>(int64_t(wd_) * cfmin_, INF - 1));
    wmax_ = std::max({wx_, wy_, int32_t(std::min<int64_t>(wd_, int64_t(wx_) + wy_))});
    vmin_ = q->via_cost >= 0 ? int32_t(std::min<int64_t>(int64_t(q->via_cost) * cfmin_, INF - 1)) : INF;
    // per-state arrays, grown as needed; epochs make a fresh search without clearing them
    const size_t NS = size_t(qc_.NS);
    if (g_.size() < NS) {
        g_.assign(NS, 0);
        seen_.assign(NS, 0);
        closed_.assign(NS, 0);
        ep_ = 0;
    }
    if (++ep_ == 0) {
        std::fill(seen_.begin(), seen_.end(), 0u);
        std::fill(closed_.begin(), closed_.end(), 0u);
        ep_ = 1;
    }
    if (hk_ == COARSE) {
        build_coarse();
    } else if (boxes && nbox) {
        boxes_.assign(boxes, boxes + nbox);
    } else {  // one box per layer around its targets
        boxes_.clear();
        std::vector<TargetBox> per(size_t(q->nl), TargetBox{-2, INT32_MAX, INT32_MAX, -1, -1});
        const int64_t words = (qc_.NS + 63) / 64;
        for (int64_t w = 0; w < words; ++w) {
            uint64_t bits = q->tbit[w];
            while (bits) {
                const int64_t ws = w * 64 + __builtin_ctzll(bits);
                bits &= bits - 1;
                if (ws >= qc_.NS) break;
                const int l = int(ws / qc_.WN);
                const int64_t c = ws % qc_.WN;
                const int i = int(c % q->ww), j = int(c / q->ww);
                TargetBox& b = per[size_t(l)];
                b.l = l;
                b.i0 = std::min(b.i0, i); b.i1 = std::max(b.i1, i);
                b.j0 = std::min(b.j0, j); b.j1 = std::max(b.j1, j);
            }
        }
        for (auto& b : per)
            if (b.l >= 0) boxes_.push_back(b);
    }
    // the starts: the least g0 of each state (as the device's k_starts)
    for (size_t k = 0; k < q->n_starts; ++k) {
        int32_t g0 = q->starts[k].value;
        const uint64_t s = q->starts[k].state;
        if (g0 >= INF || s >= uint64_t(qc_.NS)) continue;
        if (g0 < 0) g0 = 0;
        const int l = int(s / qc_.WN);
        const int64_t c = int64_t(s % qc_.WN);
        push(uint32_t(s), g0, l, int(c % q->ww), int(c / q->ww));
    }
    return true;
}

void Astar::push(uint32_t s, int32_t g, int l, int i, int j, uint8_t code, uint8_t rec) {
    if (g >= INF) return;  // the device's distances saturate there: such a state counts as unreached
    if (seen_[s] == ep_ && g >= g_[s]) return;
    const int32_t hv = h(l, i, j);
    if (hv >= INF) return;  // no target can be reached from here
    g_[s] = g;
    seen_[s] = ep_;
    if (rules_) {
        par_[s] = code;
        rec_[s] = rec;
    }
    closed_[s] = 0;  // a state reached more cheaply is searched again (only an inconsistent heuristic does that)
    heap_.push_back({g + hv, g, s});
    std::push_heap(heap_.begin(), heap_.end(), ItemLess());
}

void Astar::set_via_rules(int32_t gapc, double gap2, const uint32_t* old_vias, size_t n_old_vias) {
    rules_ = true;
    gapc_ = gapc;
    gap2_ = gap2;
    old_vias_.assign(old_vias, old_vias + (old_vias ? n_old_vias : 0));
    if (par_.size() < g_.size()) {
        par_.assign(g_.size(), 255);
        rec_.assign(g_.size(), 255);
    }
    for (const Item& it : heap_) {  // the starts begin a path: no move into them, no via yet
        par_[it.s] = 255;
        rec_[it.s] = 255;
    }
}

// a via may not sit at (i, j) (window cell): too close to one of the net's old vias other than one at the cell
bool Astar::via_blocked(int i, int j) const {
    const pcbgpu_query* q = qc_.q;
    const double x = double(q->wi0 + i), y = double(q->wj0 + j);
    const uint32_t cell = uint32_t(size_t(q->wj0 + j) * size_t(q->nx) + size_t(q->wi0 + i));
    for (uint32_t v : old_vias_) {
        const double dx = double(v % uint32_t(q->nx)) - x, dy = double(v / uint32_t(q->nx)) - y;
        if (v != cell && dx * dx + dy * dy < gap2_) return true;
    }
    return false;
}

int64_t Astar::chain(uint32_t* states, uint8_t* codes, size_t cap) const {
    if (!rules_ || best < 0) return PCBGPU_ERR_FAIL;
    const pcbgpu_query* q = qc_.q;
    const int64_t WN = qc_.WN;
    int64_t s = best, n = 0;
    for (int64_t guard = 0; guard <= qc_.NS; ++guard) {
        if (size_t(n) >= cap) return PCBGPU_ERR_FAIL;
        const uint8_t code = par_[size_t(s)];
        states[n] = uint32_t(s);
        codes[n] = code;
        ++n;
        if (code == 255) return n;
        const int l = int(s / WN);
        const int64_t c = s % WN;
        int i = int(c % q->ww), j = int(c / q->ww), pl = l;
        if (code < 8) { i -= DXS[code]; j -= DYS[code]; }
        else pl = code - 8;
        s = (int64_t(pl) * q->wh + j) * q->ww + i;
    }
    return PCBGPU_ERR_FAIL;
}

Astar::Status Astar::run(size_t budget) {
    const pcbgpu_query* q = qc_.q;
    if (!q) return BAD;
    const int64_t WN = qc_.WN;
    const int ww = q->ww, wh = q->wh, nl = q->nl;
    size_t used = 0;
    while (!heap_.empty()) {
        if (found_ && heap_.front().f > opt) break;  // every state that can tie the optimum is done
        if (budget && used >= budget) return broken_ ? BAD : BUDGET;
        std::pop_heap(heap_.begin(), heap_.end(), ItemLess());
        const Item it = heap_.back();
        heap_.pop_back();
        ++used;
        ++pops;
        const uint32_t s = it.s;
        if (it.g != g_[s] || closed_[s] == ep_) continue;  // stale
        closed_[s] = ep_;
        const int l = int(s / WN);
        const int64_t c = int64_t(s % WN);
        const int i = int(c % ww), j = int(c / ww);
        if (qc_.target(s)) {
            if (!found_ || it.g < opt || (it.g == opt && int64_t(s) < best)) {
                found_ = true;
                opt = it.g;
                best = s;
            }
            if (reverse_) return FOUND;  // the probe only asks whether a target is reachable
        }
        const int32_t g = it.g;
        if (!reverse_) {
            // forward: into the neighbours that may be entered, then vias onto the layers where one may land
            const uint8_t rq = rules_ ? rec_[s] : 255, rstep = rq == 255 ? 255 : uint8_t(std::min(254, rq + 1));
            if (q->lay[l])
                for (int dm = 0; dm < 8; ++dm) {
                    const int ni = i + DXS[dm], nj = j + DYS[dm];
                    if (ni < 0 || nj < 0 || ni >= ww || nj >= wh) continue;
                    const uint32_t w = qc_.word(l, ni, nj);
                    if (w & BLK_BIT) continue;
                    broken_ |= int32_t(w & CF_MASK) < cfmin_;
                    push(uint32_t(int64_t(l) * WN + int64_t(nj) * ww + ni), g + q->dirw[l][dm] * int32_t(w & CF_MASK), l,
                         ni, nj, uint8_t(dm), rstep);
                }
            const bool via_ok = !rules_ || ((rq == 255 || rq >= gapc_) && !via_blocked(i, j));
            if (q->via_cost >= 0 && via_ok) {
                const int32_t vc = qc_.via(i, j);
                broken_ |= vc < vmin_;
                if (vc < INF)
                    for (int l2 = 0; l2 < nl; ++l2) {
                        if (l2 == l || !(qc_.word(l2, i, j) & VOK_BIT)) continue;
                        push(uint32_t(int64_t(l2) * WN + c), g + vc, l2, i, j, uint8_t(8 + l), 0);
                    }
            }
        } else {
            // reversed: a forward move from a neighbour into this state exists only if this state may be entered
            const uint32_t w = qc_.word(l, i, j);
            broken_ |= !(w & BLK_BIT) && int32_t(w & CF_MASK) < cfmin_;
            if (q->lay[l] && !(w & BLK_BIT))
                for (int dm = 0; dm < 8; ++dm) {
                    const int pi = i - DXS[dm], pj = j - DYS[dm];
                    if (pi < 0 || pj < 0 || pi >= ww || pj >= wh) continue;
                    push(uint32_t(int64_t(l) * WN + int64_t(pj) * ww + pi), g + q->dirw[l][dm] * int32_t(w & CF_MASK), l,
                         pi, pj);
                }
            if (q->via_cost >= 0 && (w & VOK_BIT)) {
                const int32_t vc = qc_.via(i, j);
                broken_ |= vc < vmin_;
                if (vc < INF)
                    for (int l2 = 0; l2 < nl; ++l2)
                        if (l2 != l) push(uint32_t(int64_t(l2) * WN + c), g + vc, l2, i, j);
            }
        }
    }
    if (broken_) return BAD;  // the estimate may have overshot: the result is not known to be the least
    return found_ ? FOUND : EXHAUSTED;
}

}  // namespace pcb

// ---------------------------------------------------------------- C ABI

extern "C" int64_t pcbgpu_cpu_search(const pcbgpu_query* q, int heuristic, const int32_t* boxes, size_t n_boxes,
                                     int32_t cf_min, size_t max_pops, int32_t via_gapc, double via_gap2,
                                     const uint32_t* old_vias,
                                     size_t n_old_vias, int64_t* best, int32_t* opt, uint64_t* pops,
                                     uint32_t* chain_states, uint8_t* chain_codes, size_t chain_cap) {
    thread_local pcb::Astar A;
    if (best) *best = -1;
    if (opt) *opt = pcb::INF;
    if (pops) *pops = 0;
    if (heuristic != 0 && heuristic != 1) return PCBGPU_ERR_FAIL;
    std::vector<pcb::TargetBox> bx;
    for (size_t k = 0; boxes && k < n_boxes; ++k) {
        const int32_t* b = boxes + 5 * k;
        bx.push_back({b[0], b[1], b[2], b[3], b[4]});
    }
    if (!A.begin(q, heuristic ? pcb::Astar::COARSE : pcb::Astar::BOXES, bx.data(), bx.size(), false, cf_min))
        return PCBGPU_ERR_FAIL;
    const pcb::Astar::Status st = A.run(max_pops);
    if (pops) *pops = A.pops;
    if (st == pcb::Astar::BUDGET) return PCBGPU_ERR_BUDGET;
    if (st == pcb::Astar::EXHAUSTED) return 0;
    if (st != pcb::Astar::FOUND) return PCBGPU_ERR_FAIL;
    if (best) *best = A.best;
    if (opt) *opt = A.opt;
    return pcb::backtrace_chain(q, A.best, [](int64_t ws) { return A.field(ws); }, via_gapc, via_gap2, old_vias,
                                n_old_vias, chain_states, chain_codes, chain_cap);
}
