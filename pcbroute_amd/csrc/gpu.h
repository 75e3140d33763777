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
    int nx, ny, NL;                // grid width, height and layers; every map is indexed [l][j * nx + i]
    int wi0, wj0, ww, wh;          // the window searched: origin and size in cells
    uint16_t me;                   // this net's id in the owner maps (PCBGPU_FREE and PCBGPU_BLOCK are reserved)
    int32_t via_cost;              // weight of a via; negative when vias are off
    int32_t dirw[8][8];            // [layer][move]: what a step of each of the 8 moves weighs on that layer
    uint8_t lay[8];                // 1 where tracks may run on the layer
    const int32_t* cst[8];         // track cost per layer (PCBGPU_INF and above: no entry; below 0: owned copper)
    const uint16_t* st[8];         // per layer, which net owns a cell for a track centre (or free / blocked)
    const int32_t* vcst;           // via cost per cell
    const uint16_t* vst;           // which net owns a cell for a via centre
    const uint64_t* tbit;          // one bit per window state: 1 on a target
    std::vector<std::pair<uint32_t, int32_t>> own_cf;   // (window state, cost factor) where this net has copper
    std::vector<std::pair<uint32_t, int32_t>> own_vcf;  // (window cell, via factor) where this net has a via
    const uint8_t* prior = nullptr;
    int pi0 = 0, pj0 = 0, pw = 0, ph = 0, pnl = 0;
    const int* pcol = nullptr;
    const int* prow = nullptr;
    int play[8];
    int32_t mlbias[256];
    uint8_t mthr = 0;
    bool corridor = false;
    std::vector<std::pair<uint32_t, int32_t>> starts;  // (window state, cost already paid to stand there)
};

// The answer of a solve: the field over the rectangle of the window it covered; states outside read as unreachable.
struct GpuOut {
    int rx0 = 0, ry0 = 0, rw = 0, rh = 0, NL = 0, ww = 0, wh = 0;
    int64_t best = -1;                   // window state of the least-cost target, -1 if none was reached
    int32_t* d = nullptr;                // the field, [l][j * rw + i]; the library's pinned buffer, returned on destruction
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
