// backtrace.cpp - pcbgpu_backtrace: the canonical backtrace (chain.cpp) over a solved rectangle's field.
#include "backtrace.h"
#include "resolve.h"

namespace {

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
    if (!q || !out || !out->d || out->best < 0) return PCBGPU_ERR_FAIL;
    if (out->ww != q->ww || out->wh != q->wh || out->nl != q->nl) return PCBGPU_ERR_FAIL;
    return pcb::backtrace_chain(q, out->best, [out](int64_t ws) { return field_at(out, ws); }, via_gapc, via_gap2,
                                old_vias, n_old_vias, chain_states, chain_codes, chain_cap);
}
