// chain.cpp - the canonical backtrace over any field (host code shared by every back end and the CPU search).
//
// Restates src/router.cpp:1602-1679 on integer costs: from the best target, the predecessor of a state is the
// first strict minimum of field(p) + cost(p -> q) over the track moves dm = 0..7 (p = q - move) and then over the
// vias from layers 0..NL-1; the walk stops at a start whose field value is its g0. The cost factors are resolved by
// the same code the prep kernels use (resolve.h), so the walk sees the costs the device solved with.
#include <algorithm>
#include <cstdlib>
#include <unordered_map>
#include <vector>

#include "backtrace.h"
#include "query_costs.h"

namespace {

constexpr int DXS[8] = {1, -1, 0, 0, 1, 1, -1, -1};
constexpr int DYS[8] = {0, 0, 1, -1, 1, -1, 1, -1};

}  // namespace

int64_t pcb::backtrace_chain(const pcbgpu_query* q, int64_t best, const std::function<int32_t(int64_t)>& field,
                             int32_t via_gapc, double via_gap2, const uint32_t* old_vias, size_t n_old_vias,
                             uint32_t* chain_states, uint8_t* chain_codes, size_t chain_cap) {
    if (!chain_states || !chain_codes) return PCBGPU_ERR_FAIL;
    pcb::QueryCosts qc;
    if (!qc.init(q) || best < 0) return PCBGPU_ERR_FAIL;
    const int64_t WN = qc.WN, NS = qc.NS;
    std::unordered_map<uint32_t, int32_t> g0;
    for (size_t k = 0; k < q->n_starts; ++k) {
        int32_t g = q->starts[k].value;
        if (g >= pcb::INF || q->starts[k].state >= uint64_t(NS)) continue;
        if (g < 0) g = 0;
        auto it = g0.find(q->starts[k].state);
        if (it == g0.end()) g0.emplace(q->starts[k].state, g);
        else it->second = std::min(it->second, g);
    }
    int64_t n = 0;
    int64_t since_via = 1 << 20;  // track steps since the last via met on the chain
    std::vector<uint32_t> vias;
    if (best >= NS) return PCBGPU_ERR_FAIL;
    int64_t qs = best;
    for (int64_t guard = 0; guard < NS; ++guard) {
        const int l = int(qs / WN);
        const int64_t c = qs % WN;
        const int i = int(c % q->ww), j = int(c / q->ww);
        const int32_t dq = field(qs);
        if (size_t(n) >= chain_cap) return PCBGPU_ERR_FAIL;
        auto st0 = g0.find(uint32_t(qs));
        if (st0 != g0.end() && dq >= st0->second) {
            chain_states[n] = uint32_t(qs);
            chain_codes[n] = 255;
            ++n;
            break;
        }
        int64_t bc = INT64_MAX, bp = -1;
        int code = 0;
        const uint32_t w = qc.word(l, i, j);
        if (q->lay[l] && (w & pcb::CF_MASK)) {
            const int64_t cf = w & pcb::CF_MASK;
            for (int dm = 0; dm < 8; ++dm) {
                const int pi = i - DXS[dm], pj = j - DYS[dm];
                if (pi < 0 || pj < 0 || pi >= q->ww || pj >= q->wh) continue;
                const int64_t p = l * WN + int64_t(pj) * q->ww + pi;
                const int64_t cst = int64_t(field(p)) + int64_t(q->dirw[l][dm]) * cf;
                if (cst < bc) { bc = cst; bp = p; code = dm; }
            }
        }
        if (q->via_cost >= 0 && (w & pcb::VOK_BIT)) {
            const int32_t vc = qc.via(i, j);
            if (vc < pcb::INF)
                for (int l2 = 0; l2 < q->nl; ++l2) {
                    if (l2 == l) continue;
                    const int64_t p = l2 * WN + c;
                    const int64_t cst = int64_t(field(p)) + vc;
                    if (cst < bc) { bc = cst; bp = p; code = 8 + l2; }
                }
        }
        if (bp < 0 || bc > dq) return PCBGPU_ERR_CHAIN;
        chain_states[n] = uint32_t(qs);
        chain_codes[n] = uint8_t(code);
        ++n;
        if (code >= 8) {  // the CPU search keeps a path's vias apart (src/router.cpp:1661-1666)
            if (since_via < via_gapc) return PCBGPU_ERR_CHAIN;
            since_via = 0;
            vias.push_back(uint32_t(size_t(j + q->wj0) * q->nx + (i + q->wi0)));
        } else
            ++since_via;
        qs = bp;
    }
    if (n == 0 || chain_codes[n - 1] != 255) return PCBGPU_ERR_CHAIN;
    for (uint32_t vc : vias)  // ...and off the net's old vias (src/router.cpp:1670-1675)
        for (size_t k = 0; k < n_old_vias; ++k) {
            const double ddx = double(old_vias[k] % q->nx) - double(vc % q->nx);
            const double ddy = double(old_vias[k] / q->nx) - double(vc / q->nx);
            if (old_vias[k] != vc && ddx * ddx + ddy * ddy < via_gap2) return PCBGPU_ERR_CHAIN;
        }
    return n;
}
