// gpu.h - the C++ face of the MI355X search back end, in the shape of pcbroute's src/gpu.h:1-57.
//
// GpuQuery / GpuOut / gpu_available / gpu_configure / gpu_solve keep the reference's names, members and return
// codes; costs are integers (see include/pcbgpu.h): cst / vcst are int32 maps, step and dirf fold into dirw (the
// whole weight of each move), mlf becomes mlbias (an additive bias per prior value). The functions forward to the
// C ABI (libpcbgpu.so), so a router written against src/gpu.h compiles against this header with those type changes.
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/pcbgpu.h"

struct GpuQuery {
    int nx, ny, NL;                // the full grid (host arrays are [l][j * nx + i])
    int wi0, wj0, ww, wh;          // search window
    uint16_t me;                   // the net, as stored in the static grids (0: free, 0xFFFF: blocked)
    int32_t via_cost;              // < 0: no vias
    int32_t dirw[8][8];            // per layer, the weight of each of the 8 moves
    uint8_t lay[8];                // layers a track may use
    const int32_t* cst[8];         // per layer cost map: >= PCBGPU_INF blocked, negative: some net's own copper
    const uint16_t* st[8];         // static obstacles for a track centre, per layer
    const int32_t* vcst;           // the same for a via centre
    const uint16_t* vst;
    const uint64_t* tbit;          // target bits over the window states
    std::vector<std::pair<uint32_t, int32_t>> own_cf;   // window state -> cost factor of the net's own cells
    std::vector<std::pair<uint32_t, int32_t>> own_vcf;  // window cell -> via factor where the net's own vias are
    const uint8_t* prior = nullptr;
    int pi0 = 0, pj0 = 0, pw = 0, ph = 0, pnl = 0;
    const int* pcol = nullptr;
    const int* prow = nullptr;
    int play[8];
    int32_t mlbias[256];
    uint8_t mthr = 0;
    bool corridor = false;
    std::vector<std::pair<uint32_t, int32_t>> starts;  // window state, initial cost
};

// The solution: costs for the part of the window that was solved (cells outside it read as unreachable).
struct GpuOut {
    int rx0 = 0, ry0 = 0, rw = 0, rh = 0, NL = 0, ww = 0, wh = 0;
    int64_t best = -1;                   // the cheapest target reached, as a window state (-1: none)
    int32_t* d = nullptr;                // pinned host memory from a pool, given back by the destructor
    size_t cap = 0;
    GpuOut() = default;
    GpuOut(const GpuOut&) = delete;
    GpuOut& operator=(const GpuOut&) = delete;
    ~GpuOut();
    int32_t operator()(int64_t q) const;   // cost of a window state (PCBGPU_INF outside the solved rectangle)
    pcbgpu_out raw{};                     // the C-ABI result the members above mirror
};

bool gpu_available();
void gpu_configure(int slots, size_t max_states);
// sweeps >= 0, -1 on failure (or a window too big for the GPU), -3 if every slot is busy and wait is false
int gpu_solve(const GpuQuery& q, GpuOut& out, bool wait = true);
