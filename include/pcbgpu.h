/* pcbgpu.h - C ABI of the MI355X grid-search back end of pcbroute.
 *
 * An exact shortest-path solver for one search window of a multi-layer clearance grid, used when a router's A*
 * search has run past its pop threshold. It solves the same problem as pcbroute's CUDA back end
 * (src/gpu.h:1-57, src/gpu.cu:1-622) on integer costs:
 *
 *   - a track move dm (0 +x, 1 -x, 2 +y, 3 -y, 4 (+,+), 5 (+,-), 6 (-,+), 7 (-,-)) from state p into state q on
 *     layer l costs dirw[l][dm] * cf(q), where cf(q) is the resolved cost factor of q (src/gpu.cu:78-105);
 *   - a via from any layer to layer l at cell c costs via_cost * vf(c) where a via may land on l
 *     (src/gpu.cu:107-115, 331-350);
 *   - distances are int32 and saturate at PCBGPU_INF.
 *
 * Plain pointers and sizes only: every pointer in a query is borrowed for the duration of the call and copied
 * before it returns (the router's cost maps change under concurrent commits; src/router.cpp:190-195).
 *
 * Canonical result. Let D be the exact distance field of the solved rectangle, opt the least D over its target
 * states and best the lowest state index among the targets with D == opt. The returned field holds D(s) where
 * D(s) < opt, opt at best, and PCBGPU_INF everywhere else. It does not depend on the order in which the device
 * relaxes the grid, so the HIP path and a CPU solve agree bit for bit.
 */
#ifndef PCBGPU_H
#define PCBGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCBGPU_INF 0x3FFFFFFF        /* unreachable distance; a cost-map value >= it means blocked */
#define PCBGPU_MAX_LAYERS 8
#define PCBGPU_CF_MAX 0xFFFFF        /* resolved cost factors are clamped into [1, PCBGPU_CF_MAX] */
#define PCBGPU_W_MAX 1023            /* largest move weight and via cost: keeps every edge cost below 2^30 */
#define PCBGPU_FREE 0
#define PCBGPU_BLOCK 0xFFFF

/* return codes (src/gpu.h:54-57) */
#define PCBGPU_ERR_FAIL (-1)         /* failure, window too big, bad query */
#define PCBGPU_ERR_CHAIN (-2)        /* backtrace: the field does not explain a path (src/router.cpp:1654-1677) */
#define PCBGPU_ERR_BUSY (-3)         /* every slot busy and wait == 0 */

/* One entry of a state list: a window state (l * ww * wh + j * ww + i) or a window cell, and a value. */
typedef struct pcbgpu_entry {
    uint32_t state;
    int32_t value;
} pcbgpu_entry;

/* The input of one solve. Mirrors GpuQuery (src/gpu.h:10-36) with integer costs:
 *   float cst/vcst      -> int32 (>= PCBGPU_INF blocked, negative: some net's own copper, st says whose)
 *   float step, dirf    -> int32 dirw[8][8] (the whole move weight: length, direction factor)
 *   float mlf[256]      -> int32 mlbias[256], an additive cost-factor bias per prior value
 *   std::vector members -> (pointer, count) pairs */
typedef struct pcbgpu_query {
    int32_t nx, ny, nl;                  /* the full grid; host arrays are [l][j * nx + i] */
    int32_t wi0, wj0, ww, wh;            /* search window */
    uint16_t me;                         /* the net as stored in the static grids */
    int32_t via_cost;                    /* < 0: no vias */
    int32_t dirw[PCBGPU_MAX_LAYERS][8];  /* per layer, weight of each move */
    uint8_t lay[PCBGPU_MAX_LAYERS];      /* layers a track may use */
    const int32_t* cst[PCBGPU_MAX_LAYERS];
    const uint16_t* st[PCBGPU_MAX_LAYERS];
    const int32_t* vcst;
    const uint16_t* vst;
    const uint64_t* tbit;                /* target bits over the window states */
    const pcbgpu_entry* own_cf;          /* window state -> cost factor of the net's own cells */
    size_t n_own_cf;
    const pcbgpu_entry* own_vcf;         /* window cell -> via factor where the net's own vias are */
    size_t n_own_vcf;
    /* learned prior (null: none): pnl layers of ph rows of pw bytes, pixel (pi0, pj0) first; full-grid column / row
     * -> prior pixel tables; router layer -> prior layer; cost bias per prior value; corridor threshold */
    const uint8_t* prior;
    int32_t pi0, pj0, pw, ph, pnl;
    const int32_t* pcol;
    const int32_t* prow;
    int32_t play[PCBGPU_MAX_LAYERS];
    int32_t mlbias[256];
    uint8_t mthr;
    uint8_t corridor;
    const pcbgpu_entry* starts;          /* window state, initial cost g0 */
    size_t n_starts;
} pcbgpu_query;

/* The solution. Mirrors GpuOut (src/gpu.h:39-50): the rectangle of the window that was solved, the cheapest target
 * (a window state, -1: none) and the field over the rectangle ([l][j * rw + i], pinned host memory from a pool). */
typedef struct pcbgpu_out {
    int32_t rx0, ry0, rw, rh, nl, ww, wh;
    int64_t best;
    int32_t* d;
    size_t cap;
    int32_t sweeps;       /* passes of the last rectangle's solve (the field is relaxed in place: may differ by a few
                             between two solves of one query; the field, rectangle and best state do not) */
    int32_t tries;        /* rectangles solved */
    double ms_upload, ms_solve, ms_total;
    uint64_t tile_visits; /* tiles the sweep kernels processed (device work, all rectangles) */
} pcbgpu_out;

/* src/gpu.h:52 gpu_available: a device exists (the first call sets up the slots). */
int pcbgpu_available(void);
/* src/gpu.h:54 gpu_configure: concurrent solves (taken before the first use) and the most states a solve may take. */
void pcbgpu_configure(int slots, size_t max_states);
/* src/gpu.h:57 gpu_solve: sweeps >= 0, PCBGPU_ERR_FAIL, or PCBGPU_ERR_BUSY when every slot is busy and !wait. */
int pcbgpu_solve(const pcbgpu_query* q, pcbgpu_out* out, int wait);
/* GpuOut::~GpuOut (src/gpu.cu:405-407): gives the field buffer back to the pool. */
void pcbgpu_out_release(pcbgpu_out* out);
/* GpuOut::operator() (src/gpu.cu:409-416): cost of a window state, PCBGPU_INF outside the solved rectangle. */
int32_t pcbgpu_out_at(const pcbgpu_out* out, int64_t window_state);

/* The canonical backtrace (src/router.cpp:1602-1679), host code shared by every back end. From out->best, walks
 * back taking at each state the first strict minimum over the track moves dm = 0..7 and then the vias from layers
 * 0..nl-1; stops at a start whose field value is its g0. Fails (PCBGPU_ERR_CHAIN) when a state has no predecessor
 * that explains its cost, when two vias of the chain are fewer than via_gapc steps apart, or when a via lies
 * within sqrt(via_gap2) cells of one of old_vias (full-grid cells). Writes the chain target-first as (window
 * state, code: 0-7 track move, 8 + l2 via from layer l2, 255 start) and returns its length. */
int64_t pcbgpu_backtrace(const pcbgpu_query* q, const pcbgpu_out* out, int32_t via_gapc, double via_gap2,
                         const uint32_t* old_vias, size_t n_old_vias, uint32_t* chain_states, uint8_t* chain_codes,
                         size_t chain_cap);

/* ---- the CPU search (src/router.cpp:1396-1781, the `search` lambda): A* over the query's window on the costs the
 * device solves, with the device's canonical result. The heap pops states in order of cost plus an estimate that
 * never overestimates; after the first target it goes on while a state could still tie the optimum, so best is the
 * lowest target state of least cost and the chain is the one pcbgpu_backtrace walks over the device's field.
 * heuristic 0: the cheapest walk over an empty grid to the nearest target box, plus a via where the box is on another
 * layer (boxes: n_boxes rows of 5 int32 {l, i0, j0, i1, j1}, window cells, inclusive, l < 0 every layer, each target
 * inside one; NULL: one box per layer around its targets); 1: a distance map over blocks of 8 x 8 cells
 * (src/router.cpp:1401-1492). cf_min: a least cost factor of every state and via of the window that the caller
 * guarantees (1 always holds; a router whose maps never go below one uncongested cell passes that), which scales the
 * estimates; a search that meets a smaller factor fails (PCBGPU_ERR_FAIL). At most max_pops heap pops (0: no
 * limit). Returns the chain length (> 0) with best, opt
 * and the chain written as pcbgpu_backtrace writes it; 0 when no target can be reached; PCBGPU_ERR_BUDGET when the
 * pops ran out; PCBGPU_ERR_CHAIN when the via rules reject the chain; PCBGPU_ERR_FAIL on a bad query. pops gets the
 * heap pops made. Host code only: it runs without a device. */
#define PCBGPU_ERR_BUDGET (-4)
int64_t pcbgpu_cpu_search(const pcbgpu_query* q, int heuristic, const int32_t* boxes, size_t n_boxes, int32_t cf_min,
                          size_t max_pops, int32_t via_gapc, double via_gap2, const uint32_t* old_vias,
                          size_t n_old_vias, int64_t* best, int32_t* opt, uint64_t* pops, uint32_t* chain_states,
                          uint8_t* chain_codes, size_t chain_cap);

/* ---- device-resident grids (no counterpart in the reference: it copies the host maps for every solve,
 * src/gpu.cu:432-439). The router's per-class cost maps live in HBM and are kept in step by span updates; a solve
 * against a grid reads them there. */
typedef struct pcbgpu_grid pcbgpu_grid;
pcbgpu_grid* pcbgpu_grid_create(int32_t nx, int32_t ny, int32_t nl);
void pcbgpu_grid_destroy(pcbgpu_grid* g);
/* copy rows [j0, j0 + nrows) of layer l (l == nl: the via grid) from host arrays laid out like the query's */
int pcbgpu_grid_upload(pcbgpu_grid* g, int32_t l, int32_t j0, int32_t nrows, const int32_t* cst, const uint16_t* st);
/* the same for the cells [i0, i0 + w) x [j0, j0 + h) only (a commit's footprint) */
int pcbgpu_grid_upload_rect(pcbgpu_grid* g, int32_t l, int32_t i0, int32_t j0, int32_t w, int32_t h, const int32_t* cst,
                            const uint16_t* st);
/* several such rectangles at once, one staged copy and one wait (a commit's footprint on every layer and the via
 * grid): rectangle k is {l, i0, j0, w, h} = rects[5k .. 5k + 4], read from cst[k] / st[k]; 0 or -1 (nothing copied) */
int pcbgpu_grid_upload_rects(pcbgpu_grid* g, int32_t n, const int32_t* rects, const int32_t* const* cst,
                             const uint16_t* const* st);
/* a solve whose cst/st/vcst/vst come from the grid (the query's pointers for them are ignored) */
int pcbgpu_solve_grid(const pcbgpu_grid* g, const pcbgpu_query* q, pcbgpu_out* out, int wait);
/* one rectangle of the window, no widening (src/gpu.cu:422-546 solve_rect). flags: PCBGPU_NOFIELD leaves the
 * canonical field on the device (out->d stays null; best and sweeps are still returned). */
#define PCBGPU_NOFIELD 1
int pcbgpu_solve_rect_grid(const pcbgpu_grid* g, const pcbgpu_query* q, int32_t rx0, int32_t ry0, int32_t rw,
                           int32_t rh, pcbgpu_out* out, int flags);

/* ---- a solve that allocates nothing (no counterpart in the reference). For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. pcbgpu_solve_grid's
 * search and result, with every device buffer of the solve taken from one fixed scratch array per device
 * (PCBGPU_NOALLOC_SCRATCH_BYTES, held from the library's load until the process ends), uploads made straight from
 * the caller's arrays, and the canonical field written into the caller's `field` (field_cap values; ww * wh * nl
 * always suffice; NULL: no field). out gets the rectangle, best, sweeps and tries; out->d and out->cap are not
 * touched. A search whose buffers do not fit the scratch fails (PCBGPU_ERR_FAIL, "needs more memory than the no-alloc
 * scratch holds"): about 12 bytes per rectangle state plus 4 per cell, i.e. up to ~5M states. Calls in one process
 * take turns on a mutex; the scratch keeps the last call's contents. It does not take one of the slots, runs on the
 * device's null stream, and calls no hipMalloc, hipHostMalloc or hipFree (host lists of a call use the C++ heap).
 * Kernel accounting (pcbgpu_stats) does not see it. Returns sweeps >= 0 or PCBGPU_ERR_FAIL. */
#define PCBGPU_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int pcbgpu_solve_grid_noalloc(const pcbgpu_grid* g, const pcbgpu_query* q, int32_t* field, size_t field_cap,
                              pcbgpu_out* out);

/* ---- rounds: the speculative searches of one PathFinder round (src/router.cpp:2119-2242) in one sequence of
 * launches. The searches are independent; each keeps its own field, bound and widening, and its result equals what
 * pcbgpu_solve_grid returns for it alone. Submitted queries (and the grids) are borrowed until
 * pcbgpu_round_launch returns. All searches of a round share one layer count. */
typedef struct pcbgpu_round pcbgpu_round;
pcbgpu_round* pcbgpu_round_begin(void);
/* a ticket >= 0, or PCBGPU_ERR_FAIL (a query without a start, or one that does not fit the grid). The target bitmaps
 * are scanned at launch, on host threads: a search without a target fails at its pcbgpu_round_result. */
int pcbgpu_round_submit(pcbgpu_round* r, const pcbgpu_grid* g, const pcbgpu_query* q);
/* the same for one given rectangle of the window, without widening (pcbgpu_solve_rect_grid's result) */
int pcbgpu_round_submit_rect(pcbgpu_round* r, const pcbgpu_grid* g, const pcbgpu_query* q, int32_t rx0, int32_t ry0,
                             int32_t rw, int32_t rh);
/* PCBGPU_NOFIELD: the round's fields stay on the device (results keep best, rectangle and sweeps). 0 or -1. */
int pcbgpu_round_flags(pcbgpu_round* r, int flags);
/* solves every submitted search; 0 or PCBGPU_ERR_FAIL */
int pcbgpu_round_launch(pcbgpu_round* r);
/* the result of a ticket (its sweeps, or PCBGPU_ERR_FAIL); the field moves into out (release it as usual) */
int pcbgpu_round_result(pcbgpu_round* r, int ticket, pcbgpu_out* out);
void pcbgpu_round_end(pcbgpu_round* r);

/* ---- tuning and diagnostics */
/* tile rows per workgroup (16, 32, -1: by layer count) and local sweeps per tile visit; 0 keeps the current value,
 * (-1, -1) restores the defaults. Returns 0 or -1. By default routing windows (rectangles of at most 512 tiles of
 * 64 x 16 cells, whose passes are latency-bound) run 32 rows and 2 local sweeps and larger rectangles rows by layer
 * count and 2 sweeps (pcbgpu_get_tile); a setting serves every rectangle. Results never depend on it. */
int pcbgpu_set_tile(int32_t tile_rows, int32_t local_sweeps);
void pcbgpu_get_tile(int32_t nl, int32_t* tile_rows, int32_t* local_sweeps);
/* delta-stepping band over tiles: a pass visits the queued tiles whose key (the least cost change that queued them)
 * lies within delta of the least key; 0 (the default) picks 20 times the cheapest straight move weight, or no band
 * for rectangles of at most 512 tiles; -1 visits every queued tile. Results do not depend on it. 0 or -1. */
int pcbgpu_set_delta(int32_t delta);
/* ---- autotuner (no counterpart in the reference). Tile rows per wavefront, local sweeps per tile visit and the
 * delta band, picked per layer count by timing a sample round; a picked setting takes precedence over
 * pcbgpu_set_tile / pcbgpu_set_delta for rounds and solves of that layer count. Results never depend on it. */
typedef struct pcbgpu_tune {
    int32_t tile_rows;     /* 16 or 32 */
    int32_t local_sweeps;  /* 1..1024 */
    int32_t delta;         /* 0 automatic, -1 off, else the band */
    int32_t tried;         /* candidates timed (pcbgpu_autotune's output) */
    double ms;             /* the sample round's time under this setting */
} pcbgpu_tune;
/* Times the sample round (n searches, qs[i] on grids[i], one layer count) under each candidate by coordinate
 * descent (rows, then two rounds of local sweeps and delta, best of reps runs each), checks that every candidate returns the same
 * best states and rectangles, installs the fastest for that layer count and copies it to best. Returns the number
 * of candidates timed, -1 on error (the previous setting stays). Call it while no other solve runs. */
int pcbgpu_autotune(const pcbgpu_grid* const* grids, const pcbgpu_query* const* qs, int n, int reps, pcbgpu_tune* best);
/* the setting in force for a layer count: 1 if the autotuner (or tune_set/load) picked it, 0 if it is the default */
int pcbgpu_tune_get(int32_t nl, pcbgpu_tune* t);
/* install a setting for a layer count (t NULL: back to the defaults). 0 or -1. */
int pcbgpu_tune_set(int32_t nl, const pcbgpu_tune* t);
/* a text table, one line per tuned layer count: "layers tile_rows local_sweeps delta ms" */
int pcbgpu_tune_save(const char* path);
/* returns the number of settings installed, -1 if the file cannot be read */
int pcbgpu_tune_load(const char* path);

/* Kernel accounting, the counterpart of the reference's GPU_PROF timing (src/gpu.cu:491,504-516). While enabled,
 * every sweep-pass launch is bracketed by HIP events on its stream and the states it processed are counted. */
typedef struct pcbgpu_stats {
    uint64_t solves;             /* rectangles solved */
    uint64_t launches;           /* sweep-pass kernel launches (those queued past convergence included) */
    uint64_t passes;             /* passes that ran before convergence was seen */
    uint64_t tiles;              /* tile visits */
    uint64_t states;             /* states carried through a full tile relaxation (the algorithmic unit) */
    uint64_t rect_state_passes;  /* sum over rectangles of states x passes */
    double kernel_ms;            /* summed duration of the sweep-pass launches (events) */
    double solve_ms;             /* summed wall time of the rectangle solves */
    uint64_t state_sweeps;       /* states times the local sweeps run over them: the state-sweep of
                                    src/gpu.cu:493-524 (a local sweep carries every state of a tile through the eight
                                    directional relaxations and the vias), the algorithmic unit of 12 bytes */
} pcbgpu_stats;
void pcbgpu_stats_enable(int on);
void pcbgpu_stats_read(pcbgpu_stats* s, int reset);

/* last error text of the calling thread ("" if none) */
const char* pcbgpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
