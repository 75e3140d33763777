// resolve.h - how a query's maps turn into the per-state cost word the sweeps read (host and device).
//
// Restates prep_cf / prep_v (src/gpu.cu:78-115) on integer costs. One definition serves the prep kernels and the
// host backtrace, so both see exactly the same cost factors.
#pragma once
#include <stdint.h>

#include "../../include/pcbgpu.h"

#if defined(__HIPCC__)
#define PCB_HD __host__ __device__ inline
#else
#define PCB_HD inline
#endif

namespace pcb {

// The per-state word: bits 0-19 the cost factor of entering the state (0: it may not be entered), bit 20 a via may
// land there, bit 21 it is a target, bit 30 it may not be entered (OR-ed into a candidate cost, it lifts the cost
// above every bound, so a blocked cell needs no test of its own).
constexpr uint32_t CF_MASK = 0xFFFFFu;
constexpr uint32_t VOK_BIT = 1u << 20;
constexpr uint32_t TG_BIT = 1u << 21;
constexpr uint32_t BLK_BIT = 1u << 30;
constexpr int32_t INF = PCBGPU_INF;

struct ResolveParams {
    int32_t me;
    int32_t has_prior, corridor, mthr;
    int32_t pi0, pj0, pw, ph;
    int32_t via_cost;
    uint8_t lay[PCBGPU_MAX_LAYERS];
    int32_t play[PCBGPU_MAX_LAYERS];
};

PCB_HD uint32_t clamp_cf(int64_t r) { return r < 1 ? 1u : r > PCBGPU_CF_MAX ? uint32_t(PCBGPU_CF_MAX) : uint32_t(r); }

// The prior value of a state of layer l (src/gpu.cu:92-97): 255 without a prior or without a prior layer, 0 off the
// net's prior map. px / py are the state's prior pixel column and row.
PCB_HD int prior_value(const ResolveParams& p, int l, const uint8_t* prior, int32_t px, int32_t py) {
    if (!p.has_prior || !p.lay[l] || p.play[l] < 0) return 255;
    const int x = px - p.pi0, y = py - p.pj0;
    if (x < 0 || y < 0 || x >= p.pw || y >= p.ph) return 0;
    return prior[(size_t(p.play[l]) * p.ph + y) * p.pw + x];
}

// prep_cf (src/gpu.cu:78-105): map value v and static owner s of a state of layer l, its target bit and prior value
PCB_HD uint32_t resolve_word(const ResolveParams& p, const int32_t* mlbias, int l, int32_t v, uint16_t s, int tg,
                             int pv) {
    const bool sok = s == PCBGPU_FREE || int32_t(s) == p.me;
    const bool corr = p.has_prior && p.corridor;
    int64_t r = 0;
    bool blocked = false;
    if (!p.lay[l] || v >= INF) blocked = true;
    else if (v < 0) {
        if (int32_t(s) == p.me) r = -int64_t(v);
        else blocked = true;
    } else r = v;
    if (p.has_prior && p.lay[l]) {
        if (corr && pv < p.mthr && !tg) blocked = true;
        else if (!blocked) r += mlbias[pv];
    }
    const bool vok = p.lay[l] ? (sok && (!corr || tg || pv >= p.mthr)) : bool(tg);
    return (blocked ? BLK_BIT : clamp_cf(r)) | (vok ? VOK_BIT : 0u) | (tg ? TG_BIT : 0u);
}

// prep_v (src/gpu.cu:107-115): the cost of a via at a cell, INF where none may be placed
PCB_HD int32_t resolve_via(const ResolveParams& p, int32_t v, uint16_t s) {
    if (p.via_cost < 0 || v >= INF) return INF;
    int64_t r;
    if (v < 0) {
        if (int32_t(s) != p.me) return INF;
        r = -int64_t(v);
    } else r = v;
    return int32_t(int64_t(clamp_cf(r)) * p.via_cost);
}

}  // namespace pcb
