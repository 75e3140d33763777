// backtrace.cpp - the canonical backtrace over a solved field (host code shared by every back end).
//
// Restates src/router.cpp:1602-1679 on integer costs: from the best target, the predecessor of a state is the
// first strict minimum of field(p) + cost(p -> q) over the track moves dm = 0..7 (p = q - move) and then over the
// vias from layers 0..NL-1; the walk stops at a start whose field value is its g0. The cost factors are resolved by
// the same code the prep kernels use (resolve.h), so the walk sees the costs the device solved with.
#include <algorithm>
#include <cstdlib>
#include <unordered_map>
#include <vector>

#include "resolve.h"

namespace {

constexpr int DXS[8] = {1, -1, 0, 0, 1, 1, -1, -1};
constexpr int DYS[8] = {0, 0, 1, -1, 1, -1, 1, -1};

int32_t field_at(const pcbgpu_out* o, int64_t ws) {
    const int64_t WN = int64_t(o->ww) * o->wh;
    const int l = int(ws / WN);
    const int64_t c = ws % WN;
    const int i = int(c % o->ww) - o->rx0, j = int(c / o->ww) - o->ry0;
    if (ws < 0 || l >= o->nl || i < 0 || j < 0 || i >= o->rw || j >= o->rh) return pcb::INF;
    return o->d[(size_t(l) * o->rh + j) * o->rw + i];
}

}  // namespace

extern "C" int64_t pcbgpu_backtrace(const pcbgpu_query* q, const pcbgpu_out* out, int32_t via_gapc, double via_gap2,
                                    const uint32_t* old_vias, size_t n_old_vias, uint32_t* chain_states,
                                    uint8_t* chain_codes, size_t chain_cap) {
    if (!q || !out || !out->d || out->best < 0 || !chain_states || !chain_codes) return PCBGPU_ERR_FAIL;
    if (q->nl < 1 || q->nl > PCBGPU_MAX_LAYERS || !q->tbit || !q->vcst || !q->vst) return PCBGPU_ERR_FAIL;
    if (out->ww != q->ww || out->wh != q->wh || out->nl != q->nl) return PCBGPU_ERR_FAIL;
    for (int l = 0; l < q->nl; ++l)
        if (!q->cst[l] || !q->st[l]) return PCBGPU_ERR_FAIL;
    pcb::ResolveParams rp{};
    rp.me = q->me;
    rp.has_prior = q->prior != nullptr;
    rp.corridor = q->prior && q->corridor;
    rp.mthr = q->mthr;
    rp.pi0 = q->pi0; rp.pj0 = q->pj0; rp.pw = q->pw; rp.ph = q->ph;
    rp.via_cost = q->via_cost;
    for (int l = 0; l < PCBGPU_MAX_LAYERS; ++l) {
        rp.lay[l] = l < q->nl ? q->lay[l] : 0;
        rp.play[l] = l < q->nl ? q->play[l] : -1;
    }
    const int64_t WN = int64_t(q->ww) * q->wh, NS = WN * q->nl;
    std::unordered_map<uint32_t, int32_t> own, ownv, g0;
    for (size_t k = 0; k < q->n_own_cf; ++k) own[q->own_cf[k].state] = q->own_cf[k].value;   // the last entry counts
    for (size_t k = 0; k < q->n_own_vcf; ++k) ownv[q->own_vcf[k].state] = q->own_vcf[k].value;
    for (size_t k = 0; k < q->n_starts; ++k) {
        int32_t g = q->starts[k].value;
        if (g >= pcb::INF || q->starts[k].state >= uint64_t(NS)) continue;
        if (g < 0) g = 0;
        auto it = g0.find(q->starts[k].state);
        if (it == g0.end()) g0.emplace(q->starts[k].state, g);
        else it->second = std::min(it->second, g);
    }
    auto word = [&](int l, int i, int j) -> uint32_t {
        const uint32_t ws = uint32_t((int64_t(l) * q->wh + j) * q->ww + i);
        const int gi = q->wi0 + i, gj = q->wj0 + j;
        const size_t cell = size_t(gj) * q->nx + gi;
        const int tg = int((q->tbit[ws >> 6] >> (ws & 63)) & 1u);
        const int pv = pcb::prior_value(rp, l, q->prior, rp.has_prior ? q->pcol[gi] : 0, rp.has_prior ? q->prow[gj] : 0);
        auto it = own.find(ws);
        const int32_t v = it != own.end() ? it->second : q->cst[l][cell];
        return pcb::resolve_word(rp, q->mlbias, l, v, q->st[l][cell], tg, pv);
    };
    auto viacost = [&](int i, int j) -> int32_t {
        const uint32_t wc = uint32_t(int64_t(j) * q->ww + i);
        const size_t cell = size_t(q->wj0 + j) * q->nx + q->wi0 + i;
        auto it = ownv.find(wc);
        return pcb::resolve_via(rp, it != ownv.end() ? it->second : q->vcst[cell], q->vst[cell]);
    };
    int64_t n = 0;
    int64_t since_via = 1 << 20;  // track steps since the last via met on the chain
    std::vector<uint32_t> vias;
    int64_t qs = out->best;
    for (int64_t guard = 0; guard < NS; ++guard) {
        const int l = int(qs / WN);
        const int64_t c = qs % WN;
        const int i = int(c % q->ww), j = int(c / q->ww);
        const int32_t dq = field_at(out, qs);
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
        const uint32_t w = word(l, i, j);
        if (q->lay[l] && (w & pcb::CF_MASK)) {
            const int64_t cf = w & pcb::CF_MASK;
            for (int dm = 0; dm < 8; ++dm) {
                const int pi = i - DXS[dm], pj = j - DYS[dm];
                if (pi < 0 || pj < 0 || pi >= q->ww || pj >= q->wh) continue;
                const int64_t p = l * WN + int64_t(pj) * q->ww + pi;
                const int64_t cst = int64_t(field_at(out, p)) + int64_t(q->dirw[l][dm]) * cf;
                if (cst < bc) { bc = cst; bp = p; code = dm; }
            }
        }
        if (q->via_cost >= 0 && (w & pcb::VOK_BIT)) {
            const int32_t vc = viacost(i, j);
            if (vc < pcb::INF)
                for (int l2 = 0; l2 < q->nl; ++l2) {
                    if (l2 == l) continue;
                    const int64_t p = l2 * WN + c;
                    const int64_t cst = int64_t(field_at(out, p)) + vc;
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
