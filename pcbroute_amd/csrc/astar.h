// astar.h - the CPU search of the back end: A* over a query's window (src/router.cpp:1396-1781, the `search`
// lambda), on the same integer costs the device solves, with a canonical finish.
//
// The heap pops states in order of f = g + h with a consistent heuristic, so a popped state's g is its exact
// distance. After the first target is popped (its cost is opt) the search goes on until the least f in the heap
// exceeds opt: then every state on a shortest path to any target of cost opt has been popped, the target chosen is
// the lowest state index among those of cost opt, and the canonical backtrace (backtrace.h) over the popped states'
// values walks exactly the chain it walks over the device's canonical field. A search found here and one finished
// on the device therefore give the same route.
//
// A search can be run in steps of a pop budget (run) and resumed, which is how the router's gate hands an
// expensive search to the device after a number of pops (src/router.cpp:1523-1524,1690-1701).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "query_costs.h"

namespace pcb {

// A box of window cells [i0, i1] x [j0, j1] on layer l (l < 0: on every layer) holding targets.
struct TargetBox {
    int32_t l, i0, j0, i1, j1;
};

class Astar {
public:
    enum Status { FOUND = 0, EXHAUSTED = 1, BUDGET = 2, BAD = 3 };
    enum Heuristic { BOXES = 0, COARSE = 1 };

    // Sets up a search of q. BOXES: h is the cheapest walk over an empty grid to the nearest target box plus a via
    // where the box lies on another layer (boxes null: one box per layer around its targets); COARSE: a distance map
    // over 8 x 8-cell blocks from the blocks holding targets, lowered by what a state can gain inside its block
    // (src/router.cpp:1401-1492). reverse: the search goes from the query's starts to its targets over the reversed
    // moves of the query's costs (a move may leave a state only where a forward move may enter it), so it exhausts
    // exactly when no forward search from the targets could reach the starts. The router's probe uses it for that
    // answer only: its costs are those of the forward moves it reverses, not a canonical result.
    // cf_min: a lower bound of every cost factor of the window (track and via) that the caller guarantees; the
    // estimates scale with it (the router's maps never go below one uncongested cell's factor). A search that meets
    // a smaller factor ends with BAD rather than a result that might not be the least.
    bool begin(const pcbgpu_query* q, Heuristic hk, const TargetBox* boxes, size_t nbox, bool reverse = false,
               int32_t cf_min = 1);
    // The reference's via spacing inside the search (src/router.cpp:1747-1756), set after begin and before run: a
    // via may leave a state only gapc or more track steps after the last via of the path that reached it, and not
    // at a cell (full grid) closer than sqrt(gap2) to one of old_vias other than itself. The search then keeps one
    // path per state (the cheapest found, as the reference's does), so its result is a valid path, not the
    // canonical one; chain() walks it.
    void set_via_rules(int32_t gapc, double gap2, const uint32_t* old_vias, size_t n_old_vias);
    // the path found by a search with via rules: states target first, as backtrace_chain writes them (codes 0-7 the
    // track move into the state, 8 + l2 a via from layer l2, 255 the start); its length or PCBGPU_ERR_FAIL
    int64_t chain(uint32_t* states, uint8_t* codes, size_t cap) const;
    // Pops at most `budget` more states (0: no limit). FOUND: best and opt are set (canonical); EXHAUSTED: no target
    // can be reached; BUDGET: the budget ran out (run may be called again to go on).
    Status run(size_t budget);

    int64_t best = -1;    // the target reached (window state)
    int32_t opt = INF;    // its cost
    uint64_t pops = 0;    // heap pops so far, stale entries included

    // the canonical field value of a window state after FOUND: its distance if popped and below opt, opt at best,
    // INF elsewhere (a state the search did not pop cannot explain a step of the canonical chain)
    int32_t field(int64_t ws) const {
        if (ws == best) return opt;
        return closed_[size_t(ws)] == ep_ && g_[size_t(ws)] < opt ? g_[size_t(ws)] : INF;
    }
    const QueryCosts& costs() const { return qc_; }

private:
    struct Item {
        int32_t f, g;
        uint32_t s;
    };
    int32_t h(int l, int i, int j) const;
    void push(uint32_t s, int32_t g, int l, int i, int j, uint8_t code = 255, uint8_t rec = 255);
    bool via_blocked(int i, int j) const;
    void build_coarse();

    QueryCosts qc_;
    Heuristic hk_ = BOXES;
    bool reverse_ = false, found_ = false;
    std::vector<TargetBox> boxes_;
    int32_t wx_ = 1, wy_ = 1, wd_ = 1, wmax_ = 1, vmin_ = INF;  // least move costs (weight x least factor); least via cost
    int32_t cfmin_ = 1;
    bool broken_ = false;  // a factor below cfmin_ was met
    // coarse map
    static constexpr int CS = 8;
    int cw_ = 0, ch_ = 0;
    std::vector<int32_t> coarse_;
    // per window state: g, and the epochs in which g was set and the state was popped
    std::vector<int32_t> g_;
    std::vector<uint32_t> seen_, closed_;
    // with via rules: the move into each state and the track steps since its path's last via (255: none yet)
    bool rules_ = false;
    int32_t gapc_ = 0;
    double gap2_ = 0;
    std::vector<uint32_t> old_vias_;
    std::vector<uint8_t> par_, rec_;
    uint32_t ep_ = 0;
    std::vector<Item> heap_;
};

}  // namespace pcb
