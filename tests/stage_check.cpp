// stage_check.cpp - host check of the upload stage (pcbroute_amd/csrc/stage.h), built with g++ by tests/test_stage.py.
//
// Every source buffer ends exactly where a PROT_NONE guard page begins, so reading one byte past the caller's
// buffer ends the process with SIGSEGV. Checks: the bytes arrive intact for every length 0..64 and a few larger,
// the pieces are 16-byte aligned and do not overlap, a piece larger than a chunk gets a chunk of its own, and a
// reset reuses the chunks.
#include <sys/mman.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../pcbroute_amd/csrc/stage.h"

struct HeapAlloc {
    static int live;
    static void* take(size_t n) { ++live; return std::malloc(n); }
    static void release(void* p) { --live; std::free(p); }
};
int HeapAlloc::live = 0;

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) { std::fprintf(stderr, "failed: %s (line %d)\n", #c, __LINE__); return 1; } \
    } while (0)

int main() {
    const size_t pg = size_t(sysconf(_SC_PAGESIZE));
    const size_t span = 64 * pg;
    uint8_t* base = static_cast<uint8_t*>(mmap(nullptr, span + pg, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0));
    CHECK(base != MAP_FAILED);
    CHECK(mprotect(base + span, pg, PROT_NONE) == 0);
    uint8_t* guard = base + span;  // the first byte that may not be read
    for (size_t k = 0; k < span; ++k) base[k] = uint8_t(k * 131 + 7);
    {
        pcb::StageT<HeapAlloc> st;
        std::vector<std::pair<const uint8_t*, size_t>> got;
        std::vector<size_t> lens;
        for (size_t n = 0; n <= 64; ++n) lens.push_back(n);
        for (size_t n : {size_t(1000), size_t(4095), size_t(4097), size_t(65537), size_t(3 * pg + 5)}) lens.push_back(n);
        for (size_t n : lens) {
            const uint8_t* src = guard - n;  // ends at the guard page
            const uint8_t* d = static_cast<const uint8_t*>(st.put(src, n));
            CHECK(d != nullptr);
            CHECK(reinterpret_cast<uintptr_t>(d) % 16 == 0);
            CHECK(std::memcmp(d, src, n) == 0);
            got.push_back({d, n});
        }
        for (size_t a = 0; a < got.size(); ++a)  // earlier pieces were not overwritten by later ones
            CHECK(std::memcmp(got[a].first, guard - got[a].second, got[a].second) == 0);
        for (size_t a = 0; a + 1 < got.size(); ++a)
            for (size_t b = a + 1; b < got.size(); ++b) {
                const uint8_t *x0 = got[a].first, *x1 = x0 + got[a].second, *y0 = got[b].first, *y1 = y0 + got[b].second;
                CHECK(x1 <= y0 || y1 <= x0 || got[a].second == 0 || got[b].second == 0);
            }
        const int chunks = HeapAlloc::live;
        // a piece larger than the minimum chunk gets a chunk of its own
        const size_t big = pcb::StageT<HeapAlloc>::MIN_CHUNK + 3;
        std::vector<uint8_t> bigsrc(big);
        for (size_t k = 0; k < big; ++k) bigsrc[k] = uint8_t(k ^ (k >> 8));
        const uint8_t* db = static_cast<const uint8_t*>(st.put(bigsrc.data(), big));
        CHECK(db && std::memcmp(db, bigsrc.data(), big) == 0);
        CHECK(HeapAlloc::live == chunks + 1);
        // a reset reuses what is there: the same pieces again take no new chunk
        st.reset();
        for (size_t n : lens) CHECK(st.put(guard - n, n) != nullptr);
        CHECK(st.put(bigsrc.data(), big) != nullptr);
        CHECK(HeapAlloc::live == chunks + 1);
        // a null source reserves space without reading
        CHECK(st.put(nullptr, 100) != nullptr);
    }
    CHECK(HeapAlloc::live == 0);
    munmap(base, span + pg);
    std::printf("stage ok\n");
    return 0;
}
