/* pcbroute.h - C ABI of the PathFinder router that drives the MI355X search back end (pcbgpu.h).
 *
 * The caller of the back end: it restates the parts of pcbroute's router that build and consume GPU searches —
 * per-class clearance grids and occupancy stamps (src/router.cpp:383-420, 1892-1963), negotiated congestion rounds
 * with rip-up (src/router.cpp:2058-2300), the query build with starts and targets (:1525-1584), the backtrace into
 * the route (:1602-1679), read-block validation of speculative searches (:2119-2242) and the learned corridor prior
 * (:1216-1223, src/mlinfer.cpp:1414-1455). Boards come from a seeded synthetic generator (footprints SOIC-16,
 * 0805, QFP-64; nets of 2-6 pins), not from DSN files. Every search runs on the GPU; there is no CPU search.
 */
#ifndef PCBROUTE_H
#define PCBROUTE_H

#include <stddef.h>
#include <stdint.h>

#include "pcbgpu.h"
#include "pcbprior.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PCBR_MAX_CLASSES 4

/* A seeded synthetic board (sizes in mm, grid pitch in um). */
typedef struct pcbr_board_spec {
    double width_mm, height_mm, cell_um;
    int32_t layers;
    int32_t n_soic16, n_0805, n_qfp64;    /* parts placed uniformly without overlap */
    int32_t n_nets;                       /* nets of 2-6 pins (geometric, p = 0.5), pins drawn among the pads */
    int32_t n_classes;                    /* net classes 1..PCBR_MAX_CLASSES, assigned round robin */
    double track_mm[PCBR_MAX_CLASSES], clear_mm[PCBR_MAX_CLASSES], via_mm[PCBR_MAX_CLASSES];
    double edge_mm;                       /* copper to board edge */
    uint64_t seed;
    int32_t n_bga256;                     /* BGA-256 parts: 16 x 16 pads of 0.4 mm at 0.8 mm pitch, placed first; a
                                             net with a BGA pin takes class 0 (the finest), others round robin */
    int32_t n_planes;                     /* inner layers reserved as planes (the middle ones): no class's tracks
                                             use them, vias pass them */
} pcbr_board_spec;

typedef struct pcbr_board pcbr_board;
pcbr_board* pcbr_board_generate(const pcbr_board_spec* spec);
void pcbr_board_free(pcbr_board* b);
/* A board from a Specctra DSN file (src/dsn.cpp:75-384, read_dsn) on a grid of cell_um cells: the signal layers, the
 * boundary (its bounding box is the grid; a polygon outline blocks what lies outside it), the default rule and the
 * classes' width, clearance and via, every part's pins placed as src/dsn.cpp:239-252 places them, each pad as the
 * bounding rectangle of its padstack's shapes, the keepouts of the structure and of the parts' images (placed with
 * their part) as rectangles, the copper-to-edge clearance of (pcbroute_rules (edge_clearance UM)), else 500 um.
 * Planes and pre-routed wiring are not read. NULL on error (pcbr_last_error). */
pcbr_board* pcbr_board_read_dsn(const char* path, double cell_um);
/* A board (generated or read) as a DSN file any DSN router reads: every pad a one-pin part on a rectangle padstack,
 * the classes with their rules and vias. 0 or -1. */
int pcbr_board_write_dsn(const pcbr_board* b, const char* path);

typedef struct pcbr_board_info {
    int32_t nx, ny, nl, n_pads, n_nets, n_pins, n_connections, n_parts;
    uint64_t hash;  /* FNV-1a over the pads and nets: the same seed gives the same board */
} pcbr_board_info;
int pcbr_board_get_info(const pcbr_board* b, pcbr_board_info* info);
/* pads as rows of 6 int32: x0, y0, x1, y1 (cells, half open), layer (-1: every layer), net (-1: none) */
int pcbr_board_pads(const pcbr_board* b, int32_t* out, size_t cap);

typedef struct pcbr_options {
    int32_t max_iters;        /* negotiation rounds before legalizing (src/router.h:38) */
    int32_t patience;         /* stop after this many rounds without a 10 % improvement */
    double pres0, pres_growth, hist_inc;
    double window_mm;         /* search window margin around a net's pins (src/router.h:15) */
    double via_cost_mm;       /* a via costs as much as this much straight track (src/router.h:16) */
    double wrong_way;         /* cost factor against a layer's preferred direction (src/router.h:24) */
    int32_t wave;             /* 0: one net at a time; > 0: up to this many nets routed together, committed in
                                 order with read-block validation (the result equals one at a time); < 0: each
                                 group of the order (group_mm) routed together, at most -wave nets at once */
    double group_mm;          /* the routing order: maximal groups of nets whose pin boxes, grown by this, do not
                                 overlap, group after group (src/router.cpp:2128-2137); < 0: each box grown by
                                 -group_mm times its longer side plus 1 mm */
    /* learned prior: applied in the first round when a network is attached (pcbr_set_prior) */
    double ml_weight;         /* a step costs up to 1 + this where the prior says no (src/router.h:59) */
    double ml_sat;            /* prior at or above this: no extra cost (src/router.h:60) */
    double ml_corridor;       /* search only where the prior is at least this, falling back to the plain search */
    double ml_margin_mm;      /* the learned window: the corridor's bounding box plus this */
    int32_t legalize;         /* after negotiation, reroute what still conflicts with others' copper as obstacles */
    int32_t verbose;
    int32_t negotiate;        /* 0: waves commit in order with validation (the routes equal one net at a time);
                                 1: from the second negotiation round on, the nets of a batch (|wave| at once) are
                                 routed against the board as the batch began and all committed, overlaps left to the
                                 next round (parallel PathFinder); 2: also in the first round; 3: as 1 until two
                                 rounds in a row bring no 10 % fewer conflicts, then in order as 0 (with a fresh
                                 patience). Legalizing is always one net at a time. */
    int32_t rescue;           /* rounds of the rescue pass after legalizing (src/router.cpp:2401-2454; 1, the default;
                                 0: none): a net still open is routed again pushing through others' copper, the
                                 nets it then crosses route around it, and the board is kept only if fewer
                                 connections are open in all, else put back */
    int32_t smooth;           /* pcbr_write_ses smooths the tracks (1, the default, as src/main.cpp:287-290) or writes
                                 the grid paths' straight runs (0) */
    int32_t ml_rounds;        /* the learned prior (network or file) guides a net's searches in the first negotiation
                                 round only (0, the default) or in every round (1, as src/router.cpp:674,1218 keep
                                 a net's map for all its searches); never while legalizing */
    /* The GPU gate (src/router.cpp:1523-1524,1690-1701). A search runs on the CPU (pcbgpu_cpu_search's A*) and is
     * finished on the device once it has made gpu_pops heap pops; 1 (the default): every search goes to the device
     * at once; 0: never (CPU only, up to max_pops). Both give the same route: the CPU search's result is the device's
     * canonical one. */
    int64_t gpu_pops;
    int64_t max_pops;         /* a CPU search's budget without the device (src/router.h:16: 3,000,000); a connection
                                 it cannot finish stays open */
    int64_t probe_pops;       /* before searching for the last open pin of a net, a search from that pin back towards
                                 the net's copper within this many pops (src/router.cpp:1785-1800; 0: none): a pin
                                 walled in is left open without a full search */
    int32_t gpu_wait;         /* wait for a free device slot (1, the default) or, when all are busy, go on on the
                                 CPU for 200,000 pops and try again (0, --gpu-nowait) */
    double time_limit_s;      /* a route's time limit (src/router.cpp:2247-2256; 0: none): negotiation stops when the
                                 next round would pass half of it, legalizing follows with rip-ups until 80 % of it,
                                 the rescue pass gets what is left */
    double rescue_s;          /* the rescue pass's own cap (0: half the time the route took so far, at least 10 s) */
    int32_t ripups;           /* while legalizing, times a net left open may push its way through others' copper, the
                                 nets it crosses then routing around it (src/router.h:42, src/router.cpp:2296-2340;
                                 3, the default; 0: none) */
} pcbr_options;
void pcbr_default_options(pcbr_options* o);

typedef struct pcbr_result {
    int32_t nets, connections, opens, violations, conflict_nets, iterations;
    int64_t searches, search_fallbacks, rounds, revalidated, length_cells, vias;
    uint64_t tile_visits;
    double route_s, search_s, prior_s, upload_s, backtrace_s;
    double host_s[4];         /* host time: cost rebuilds, stamps, read sets, query builds */
    uint64_t route_hash;      /* FNV-1a over every net's routed cells in order: equal routes, equal hash */
    /* where the searches ended (the gate): found or exhausted on the CPU, finished on the device, device failures
     * that the CPU finished, chains the via spacing rejected (searched again on the CPU under the rules), pins the
     * probe found walled in, connections left open when a CPU budget ran out; heap pops and time on the CPU */
    int64_t cpu_searches, gpu_searches, gpu_failures, chain_rejects, probe_opens, budget_opens;
    uint64_t cpu_pops;
    double cpu_s;
    int32_t first_conflict_nets, first_conflict_cells;  /* left in conflict by the first negotiation round */
    int32_t rescue_tried, rescue_kept, opens_before_rescue;
    int32_t limited;          /* 1: the time limit cut negotiation short, 2: also the rescue pass */
    int32_t ripups;           /* pushes made while legalizing */
    uint64_t solved_states;   /* states the searches covered: each device search's solved rectangle, each CPU
                                 search's window (the work a search costs, whichever side ran it) */
} pcbr_result;

typedef struct pcbr_router pcbr_router;
pcbr_router* pcbr_create(const pcbr_board* b, const pcbr_options* o);
void pcbr_destroy(pcbr_router* r);
/* routes the board; 0 or -1 (pcbr_last_error) */
int pcbr_route(pcbr_router* r, pcbr_result* res);
/* the independent check of the committed routes (the counterpart of --eval and check_routes,
 * src/eval.cpp:67-220, src/router.cpp:2864-2936): copper connectivity per net and clearance between nets'
 * tracks, vias and pads, recomputed from the routes alone */
int pcbr_eval(pcbr_router* r, int32_t* opens, int32_t* violations);
/* a net's route as rows of 3 int32 (layer, i, j) in path order, paths separated by a row (-1, -1, -1) */
int64_t pcbr_net_route(const pcbr_router* r, int32_t net, int32_t* out, size_t cap_rows);

/* Called after every search with its query and canonical result (field included), for checking. */
typedef void (*pcbr_search_cb)(void* user, const pcbgpu_query* q, const pcbgpu_out* out);
void pcbr_set_search_callback(pcbr_router* r, pcbr_search_cb cb, void* user);

/* ---- the learned corridor prior (replaces the RouteNet call, src/mlinfer.cpp:834-856) */
#define PCBR_FEAT_PER_LAYER 5  /* obstacles, others' occupancy, history, this net's pads, other nets' pads */
#define PCBR_FEAT_NET 2        /* then, once per pixel: closeness to the net's pin spanning tree, its pins' box */
#define PCBR_PIX 4             /* a prior pixel is 4 x 4 grid cells */
/* attach a network (in_ch = PCBR_FEAT_PER_LAYER * layers + PCBR_FEAT_NET, out_ch = layers); NULL detaches */
int pcbr_set_prior(pcbr_router* r, pcbprior* net);
/* The prior file (MLP1; read src/mlprior.cpp:46-74, write src/mlinfer.cpp:1556-1577): "MLP1", f64 pitch, x0, y0
 * (micrometres), u32 layer count and per layer its name (u32 length, bytes), u32 net count, then per net its name,
 * i32 i0, j0, u32 w, h and w * h * layers bytes (layer-major, then rows, then columns) over pixels [i0, i0 + w) x
 * [j0, j0 + h). Here net k is named "N<k>", layer l "L<l>", the board's corner is (0, 0) and a pixel is PCBR_PIX
 * cells. pcbr_write_prior writes the priors the network gave the last route's first round (nets in name order);
 * returns the net count or -1. pcbr_read_prior loads a file whose maps then stand in for a network (while none is
 * attached) in the first round of each route: cells map to the file's pixels and layers match by name as in
 * src/router.cpp:660-687 (outside a net's map 0, a layer the file lacks no say); nets without a map route plainly.
 * Returns the number of the board's nets the file has a map for, or -1 (bad magic, pitch, sizes, a short file). */
int pcbr_write_prior(const pcbr_router* r, const char* path);
/* The routes as a Specctra session file (write_ses, src/router.cpp:2938-2963): per net its wires at its class's width
 * and a via of its class at every layer change; the board's names and coordinates (a generated board: nets N<k>,
 * layers L<l>, vias V<class>). With the smooth option (smooth_routes, src/smooth.cpp:105-225) each chain of track
 * between fixed vertices (vias, the net's pads, ends, junctions) is redrawn with the fewest 0/45/90-degree segments
 * that clear other nets' copper, keepouts and the board edge exactly, nets one after another; otherwise the wires are
 * the grid paths' straight runs from cell centre to cell centre. 0 or -1. */
int pcbr_write_ses(const pcbr_router* r, const char* path);
int pcbr_read_prior(pcbr_router* r, const char* path);
/* Training samples: when enabled, the first route of every net in the first round records its input features
 * (the board as that net's search saw it) over a crop around its pins; after pcbr_route, the net's final route is
 * the target. Crops are ch x cw pixels (cw a multiple of 64), placed around the pins' box and inside the board. */
int pcbr_record_samples(pcbr_router* r, int32_t ch, int32_t cw);
int32_t pcbr_num_samples(const pcbr_router* r);
/* sample k: x (ch, cw, in_ch), target (ch, cw, layers) in {0, 1}, mask (ch, cw) */
int pcbr_get_sample(const pcbr_router* r, int32_t k, float* x, float* target, float* mask);

const char* pcbr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
