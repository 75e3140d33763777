// gpu_shim.cpp - src/gpu.h's functions over the C ABI (see gpu.h beside this file).
#include "gpu.h"

#include <cstring>

namespace {

static_assert(sizeof(std::pair<uint32_t, int32_t>) == sizeof(pcbgpu_entry), "a list entry is (state, value)");

const pcbgpu_entry* entries(const std::vector<std::pair<uint32_t, int32_t>>& v) {
    return v.empty() ? nullptr : reinterpret_cast<const pcbgpu_entry*>(v.data());
}

}  // namespace

GpuOut::~GpuOut() { pcbgpu_out_release(&raw); }

int32_t GpuOut::operator()(int64_t q) const { return pcbgpu_out_at(&raw, q); }

bool gpu_available() { return pcbgpu_available() != 0; }

void gpu_configure(int slots, size_t max_states) { pcbgpu_configure(slots, max_states); }

int gpu_solve(const GpuQuery& q, GpuOut& out, bool wait) {
    pcbgpu_query c{};
    c.nx = q.nx; c.ny = q.ny; c.nl = q.NL;
    c.wi0 = q.wi0; c.wj0 = q.wj0; c.ww = q.ww; c.wh = q.wh;
    c.me = q.me;
    c.via_cost = q.via_cost;
    std::memcpy(c.dirw, q.dirw, sizeof c.dirw);
    std::memcpy(c.lay, q.lay, sizeof c.lay);
    for (int l = 0; l < PCBGPU_MAX_LAYERS; ++l) {
        c.cst[l] = l < q.NL ? q.cst[l] : nullptr;
        c.st[l] = l < q.NL ? q.st[l] : nullptr;
        c.play[l] = q.play[l];
    }
    c.vcst = q.vcst; c.vst = q.vst; c.tbit = q.tbit;
    c.own_cf = entries(q.own_cf); c.n_own_cf = q.own_cf.size();
    c.own_vcf = entries(q.own_vcf); c.n_own_vcf = q.own_vcf.size();
    c.prior = q.prior; c.pi0 = q.pi0; c.pj0 = q.pj0; c.pw = q.pw; c.ph = q.ph; c.pnl = q.pnl;
    c.pcol = q.pcol; c.prow = q.prow;
    std::memcpy(c.mlbias, q.mlbias, sizeof c.mlbias);
    c.mthr = q.mthr;
    c.corridor = q.corridor ? 1 : 0;
    c.starts = entries(q.starts); c.n_starts = q.starts.size();
    const int rc = pcbgpu_solve(&c, &out.raw, wait ? 1 : 0);
    out.rx0 = out.raw.rx0; out.ry0 = out.raw.ry0; out.rw = out.raw.rw; out.rh = out.raw.rh;
    out.NL = out.raw.nl; out.ww = out.raw.ww; out.wh = out.raw.wh;
    out.best = out.raw.best;
    out.d = out.raw.d;
    out.cap = out.raw.cap;
    return rc;
}
