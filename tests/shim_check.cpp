// shim_check.cpp - solves a seeded synthetic search through the C++ gpu.h shim and through the C ABI and checks that
// both give the same rectangle, best target and field. Prints "ok" and exits 0 when they agree.
#include <cstdio>
#include <cstring>
#include <random>

#include "../pcbroute_amd/csrc/gpu.h"

int main() {
    if (!gpu_available()) { std::printf("no device\n"); return 3; }
    const int nx = 300, ny = 220, NL = 3;
    std::mt19937 rng(5);
    std::vector<int32_t> cst(size_t(NL) * nx * ny), vcst(size_t(nx) * ny);
    std::vector<uint16_t> st(size_t(NL) * nx * ny, 0), vst(size_t(nx) * ny, 0);
    for (auto& c : cst) c = (rng() % 10 == 0) ? PCBGPU_INF : int32_t(1 + rng() % 6);
    for (size_t k = 0; k < cst.size(); ++k)
        if (cst[k] >= PCBGPU_INF) st[k] = PCBGPU_BLOCK;
    for (auto& c : vcst) c = int32_t(1 + rng() % 3);
    GpuQuery q{};
    q.nx = nx; q.ny = ny; q.NL = NL;
    q.wi0 = 10; q.wj0 = 7; q.ww = 250; q.wh = 190;
    q.me = 0;
    q.via_cost = 300;
    for (int l = 0; l < 8; ++l) {
        for (int d = 0; d < 8; ++d) q.dirw[l][d] = d < 4 ? 12 + l : 17 + l;
        q.lay[l] = l < NL;
        q.play[l] = -1;
        if (l < NL) { q.cst[l] = cst.data() + size_t(l) * nx * ny; q.st[l] = st.data() + size_t(l) * nx * ny; }
    }
    q.vcst = vcst.data(); q.vst = vst.data();
    for (auto& b : q.mlbias) b = 0;
    const size_t WN = size_t(q.ww) * q.wh, WS = WN * NL;
    std::vector<uint64_t> tb(WS / 64 + 1, 0);
    const uint32_t t = uint32_t(2 * WN + 170 * q.ww + 230);
    tb[t >> 6] |= 1ull << (t & 63);
    q.tbit = tb.data();
    q.starts = {{uint32_t(0 * WN + 12 * q.ww + 9), 0}, {uint32_t(1 * WN + 14 * q.ww + 11), 24}};
    q.own_cf = {{uint32_t(0 * WN + 13 * q.ww + 10), 1}};
    GpuOut a;
    const int ra = gpu_solve(q, a);
    pcbgpu_query c{};
    c.nx = nx; c.ny = ny; c.nl = NL; c.wi0 = q.wi0; c.wj0 = q.wj0; c.ww = q.ww; c.wh = q.wh; c.me = 0; c.via_cost = 300;
    std::memcpy(c.dirw, q.dirw, sizeof c.dirw);
    std::memcpy(c.lay, q.lay, sizeof c.lay);
    for (int l = 0; l < NL; ++l) { c.cst[l] = q.cst[l]; c.st[l] = q.st[l]; }
    for (int l = 0; l < 8; ++l) c.play[l] = -1;
    c.vcst = q.vcst; c.vst = q.vst; c.tbit = q.tbit;
    pcbgpu_entry s0[2] = {{q.starts[0].first, q.starts[0].second}, {q.starts[1].first, q.starts[1].second}};
    pcbgpu_entry o0[1] = {{q.own_cf[0].first, q.own_cf[0].second}};
    c.starts = s0; c.n_starts = 2; c.own_cf = o0; c.n_own_cf = 1;
    pcbgpu_out b{};
    const int rb = pcbgpu_solve(&c, &b, 1);
    if (ra < 0 || rb < 0) { std::printf("solve failed %d %d: %s\n", ra, rb, pcbgpu_last_error()); return 1; }
    bool same = a.rx0 == b.rx0 && a.ry0 == b.ry0 && a.rw == b.rw && a.rh == b.rh && a.best == b.best && ra == rb;
    const size_t n = size_t(a.rw) * a.rh * NL;
    for (size_t k = 0; same && k < n; ++k) same = a.d[k] == b.d[k];
    const int32_t cost = a(a.best);
    pcbgpu_out_release(&b);
    std::printf("%s best %lld cost %d sweeps %d\n", same ? "ok" : "DIFFER", (long long)a.best, cost, ra);
    return same ? 0 : 2;
}
