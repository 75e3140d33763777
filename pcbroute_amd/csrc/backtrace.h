// backtrace.h - the canonical backtrace over any field (host code shared by every back end and the CPU search).
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>

#include "../../include/pcbgpu.h"

namespace pcb {

// The chain from `best` back to a start (src/router.cpp:1602-1679) over the field value of each window state that
// `field` gives (PCBGPU_INF where the search holds none). Same rules and return values as pcbgpu_backtrace.
int64_t backtrace_chain(const pcbgpu_query* q, int64_t best, const std::function<int32_t(int64_t)>& field,
                        int32_t via_gapc, double via_gap2, const uint32_t* old_vias, size_t n_old_vias,
                        uint32_t* chain_states, uint8_t* chain_codes, size_t chain_cap);

}  // namespace pcb
