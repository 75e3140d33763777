// route_on_oracle.cpp - the router of pcbroute_amd/csrc/router.cpp over the CPU oracle's search.
//
// TEST INFRASTRUCTURE ONLY: built by the tests into oracle/libroute_oracle.so together with router.cpp. It serves the
// pcbgpu_* calls the router makes (device-resident grids, solves, rounds, backtrace) from host copies of the maps
// and the oracle's exact solve (pcb_oracle.c), so the router's host logic runs and is checked without a GPU, and its
// routes can be compared with those the HIP back end gives. The product library never links this file.
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/pcbgpu.h"
#include "../include/pcbprior.h"
#include "pcb_oracle.h"

struct pcbgpu_grid {
    int nx, ny, nl;
    std::vector<int32_t> cst, vcst;
    std::vector<uint16_t> st, vst;
};

struct pcbgpu_round {
    std::vector<std::pair<const pcbgpu_grid*, pcbgpu_query>> q;
    std::vector<pcbgpu_out> out;
    std::vector<int> rc;
};

namespace {
std::string g_err;
// failures planted by the tests (route_oracle_inject): every k-th solve fails, every k-th backtrace is rejected
int g_fail_every = 0, g_chain_every = 0;
long g_solves = 0, g_traces = 0;
bool planted_failure() { return g_fail_every > 0 && ++g_solves % g_fail_every == 0; }

pcbgpu_query on_grid(const pcbgpu_grid* g, const pcbgpu_query* q) {
    pcbgpu_query x = *q;
    const size_t n = size_t(g->nx) * g->ny;
    for (int l = 0; l < g->nl; ++l) {
        x.cst[l] = g->cst.data() + size_t(l) * n;
        x.st[l] = g->st.data() + size_t(l) * n;
    }
    x.vcst = g->vcst.data();
    x.vst = g->vst.data();
    return x;
}
}  // namespace

extern "C" {

pcbgpu_grid* pcbgpu_grid_create(int32_t nx, int32_t ny, int32_t nl) {
    auto* g = new pcbgpu_grid{nx, ny, nl, {}, {}, {}, {}};
    const size_t n = size_t(nx) * ny;
    g->cst.assign(n * nl, PCBGPU_INF);
    g->st.assign(n * nl, PCBGPU_BLOCK);
    g->vcst.assign(n, PCBGPU_INF);
    g->vst.assign(n, PCBGPU_BLOCK);
    return g;
}

void pcbgpu_grid_destroy(pcbgpu_grid* g) { delete g; }

int pcbgpu_grid_upload(pcbgpu_grid* g, int32_t l, int32_t j0, int32_t nrows, const int32_t* cst, const uint16_t* st) {
    if (!g || l < 0 || l > g->nl || j0 < 0 || nrows < 0 || j0 + nrows > g->ny) return -1;
    const size_t n = size_t(g->nx) * g->ny, off = size_t(j0) * g->nx, cnt = size_t(nrows) * g->nx;
    int32_t* dc = l < g->nl ? g->cst.data() + size_t(l) * n : g->vcst.data();
    uint16_t* ds = l < g->nl ? g->st.data() + size_t(l) * n : g->vst.data();
    std::memcpy(dc + off, cst + off, cnt * 4);
    std::memcpy(ds + off, st + off, cnt * 2);
    return 0;
}

int pcbgpu_grid_upload_rect(pcbgpu_grid* g, int32_t l, int32_t i0, int32_t j0, int32_t w, int32_t h, const int32_t* cst,
                            const uint16_t* st) {
    if (!g || l < 0 || l > g->nl || i0 < 0 || j0 < 0 || w < 0 || h < 0 || i0 + w > g->nx || j0 + h > g->ny) return -1;
    const size_t n = size_t(g->nx) * g->ny;
    int32_t* dc = l < g->nl ? g->cst.data() + size_t(l) * n : g->vcst.data();
    uint16_t* ds = l < g->nl ? g->st.data() + size_t(l) * n : g->vst.data();
    for (int j = j0; j < j0 + h; ++j) {
        const size_t off = size_t(j) * g->nx + i0;
        std::memcpy(dc + off, cst + off, size_t(w) * 4);
        std::memcpy(ds + off, st + off, size_t(w) * 2);
    }
    return 0;
}

int pcbgpu_grid_upload_rects(pcbgpu_grid* g, int32_t n, const int32_t* rects, const int32_t* const* cst,
                             const uint16_t* const* st) {
    for (int k = 0; k < n; ++k)
        if (pcbgpu_grid_upload_rect(g, rects[5 * k], rects[5 * k + 1], rects[5 * k + 2], rects[5 * k + 3],
                                    rects[5 * k + 4], cst[k], st[k]) != 0)
            return -1;
    return 0;
}

void route_oracle_inject(int fail_every, int chain_every) {
    g_fail_every = fail_every;
    g_chain_every = chain_every;
    g_solves = g_traces = 0;
}

int pcbgpu_solve_grid(const pcbgpu_grid* g, const pcbgpu_query* q, pcbgpu_out* out, int) {
    if (planted_failure()) { *out = pcbgpu_out{}; g_err = "planted failure"; return PCBGPU_ERR_FAIL; }
    const pcbgpu_query x = on_grid(g, q);
    return orc_solve(&x, out, size_t(1) << 31);
}

void pcbgpu_out_release(pcbgpu_out* out) {
    if (out && out->d) orc_out_free(out);
}

pcbgpu_round* pcbgpu_round_begin(void) { return new pcbgpu_round; }

int pcbgpu_round_submit(pcbgpu_round* r, const pcbgpu_grid* g, const pcbgpu_query* q) {
    r->q.push_back({g, *q});
    return int(r->q.size()) - 1;
}

int pcbgpu_round_launch(pcbgpu_round* r) {
    r->out.assign(r->q.size(), pcbgpu_out{});
    r->rc.assign(r->q.size(), -1);
    for (size_t k = 0; k < r->q.size(); ++k) {
        const pcbgpu_query x = on_grid(r->q[k].first, &r->q[k].second);
        r->rc[k] = planted_failure() ? PCBGPU_ERR_FAIL : orc_solve(&x, &r->out[k], size_t(1) << 31);
    }
    return 0;
}

int pcbgpu_round_result(pcbgpu_round* r, int t, pcbgpu_out* out) {
    *out = r->out[size_t(t)];
    r->out[size_t(t)].d = nullptr;
    return r->rc[size_t(t)];
}

void pcbgpu_round_end(pcbgpu_round* r) {
    for (auto& o : r->out) pcbgpu_out_release(&o);
    delete r;
}

int64_t pcbgpu_backtrace(const pcbgpu_query* q, const pcbgpu_out* out, int32_t via_gapc, double via_gap2,
                         const uint32_t* old_vias, size_t n_old_vias, uint32_t* chain_states, uint8_t* chain_codes,
                         size_t chain_cap) {
    if (g_chain_every > 0 && ++g_traces % g_chain_every == 0) return PCBGPU_ERR_CHAIN;
    return orc_backtrace(q, out, via_gapc, via_gap2, old_vias, n_old_vias, chain_states, chain_codes, chain_cap);
}

const char* pcbgpu_last_error(void) { return g_err.c_str(); }

int pcbprior_predict(pcbprior*, const float*, int, int, int, float*) { return -1; }
int pcbprior_predict_valid(pcbprior*, const float*, int, int, int, const int32_t*, float*) { return -1; }
const char* pcbprior_last_error(void) { return "no network in the oracle build"; }

}  // extern "C"
