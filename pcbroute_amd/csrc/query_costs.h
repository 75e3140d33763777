// query_costs.h - a query's costs as the host sees them: the cost word of a window state and the via cost of a
// window cell, resolved by the code the prep kernels run (resolve.h), with the query's own lists applied. The host
// backtrace and the CPU search read the costs through it, so every back end relaxes and walks the same costs.
#pragma once
#include <cstdint>
#include <unordered_map>

#include "resolve.h"

namespace pcb {

struct QueryCosts {
    const pcbgpu_query* q = nullptr;
    ResolveParams rp{};
    int64_t WN = 0, NS = 0;
    std::unordered_map<uint32_t, int32_t> own, ownv;  // window state / cell -> value (the last entry of a list counts)

    // false on a query the host cannot read (null maps, bad layer count)
    bool init(const pcbgpu_query* query) {
        q = query;
        if (!q || q->nl < 1 || q->nl > PCBGPU_MAX_LAYERS || !q->tbit || !q->vcst || !q->vst) return false;
        if (q->ww < 1 || q->wh < 1) return false;
        for (int l = 0; l < q->nl; ++l)
            if (!q->cst[l] || !q->st[l]) return false;
        if (q->prior && (!q->pcol || !q->prow)) return false;
        rp = ResolveParams{};
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
        WN = int64_t(q->ww) * q->wh;
        NS = WN * q->nl;
        own.clear();
        ownv.clear();
        for (size_t k = 0; k < q->n_own_cf; ++k) own[q->own_cf[k].state] = q->own_cf[k].value;
        for (size_t k = 0; k < q->n_own_vcf; ++k) ownv[q->own_vcf[k].state] = q->own_vcf[k].value;
        return true;
    }
    bool target(int64_t ws) const { return (q->tbit[ws >> 6] >> (ws & 63)) & 1u; }
    // the cost word of window state (l, i, j): cost factor, via landing, target and blocked bits
    uint32_t word(int l, int i, int j) const {
        const int64_t ws = (int64_t(l) * q->wh + j) * q->ww + i;
        const int gi = q->wi0 + i, gj = q->wj0 + j;
        const size_t cell = size_t(gj) * size_t(q->nx) + size_t(gi);
        const int pv = prior_value(rp, l, q->prior, rp.has_prior ? q->pcol[gi] : 0, rp.has_prior ? q->prow[gj] : 0);
        int32_t v = q->cst[l][cell];
        if (!own.empty()) {
            auto it = own.find(uint32_t(ws));
            if (it != own.end()) v = it->second;
        }
        return resolve_word(rp, q->mlbias, l, v, q->st[l][cell], target(ws) ? 1 : 0, pv);
    }
    // the cost of a via at window cell (i, j), INF where none may be placed
    int32_t via(int i, int j) const {
        const size_t cell = size_t(q->wj0 + j) * size_t(q->nx) + size_t(q->wi0 + i);
        int32_t v = q->vcst[cell];
        if (!ownv.empty()) {
            auto it = ownv.find(uint32_t(int64_t(j) * q->ww + i));
            if (it != ownv.end()) v = it->second;
        }
        return resolve_via(rp, v, q->vst[cell]);
    }
};

}  // namespace pcb
