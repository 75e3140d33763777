// stage.h - host staging for the small uploads of one search (lists, tables, target words).
//
// A copy from pinned memory is truly asynchronous, so a solve queues its uploads without waiting for them. The
// stage hands out 16-byte aligned pieces of a few large chunks; it is reset for each query and keeps its chunks for
// the next. The allocator is a policy (pinned host memory in the library, plain heap memory in the host test), so
// the bookkeeping is checked on the CPU.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace pcb {

template <class Alloc>
struct StageT {
    static constexpr size_t ALIGN = 16;
    static constexpr size_t MIN_CHUNK = size_t(1) << 20;
    std::vector<std::pair<uint8_t*, size_t>> chunks;
    size_t cur = 0, used = 0;
    StageT() = default;
    StageT(const StageT&) = delete;
    StageT& operator=(const StageT&) = delete;
    ~StageT() { for (auto& c : chunks) Alloc::release(c.first); }
    void reset() { cur = 0; used = 0; }
    // Copies n bytes of src (if not null) into the stage and returns where they are; null when no memory. Exactly n
    // bytes are read from src: the piece is rounded up to 16 bytes only in the stage's own bookkeeping, so a source
    // that ends at the end of a mapping is never read past.
    const void* put(const void* src, size_t n) {
        const size_t span = (n + ALIGN - 1) & ~(ALIGN - 1);
        while (cur < chunks.size() && used + span > chunks[cur].second) { ++cur; used = 0; }
        if (cur == chunks.size()) {
            const size_t cap = std::max(span, MIN_CHUNK);
            uint8_t* p = static_cast<uint8_t*>(Alloc::take(cap));
            if (!p) return nullptr;
            chunks.push_back({p, cap});
            used = 0;
        }
        uint8_t* d = chunks[cur].first + used;
        if (src && n) std::memcpy(d, src, n);
        used += span;
        return d;
    }
};

}  // namespace pcb
