/* pcb_oracle.h - CPU restatement of pcbroute's grid-search back end on integer costs.
 *
 * TEST INFRASTRUCTURE ONLY. The product (libpcbgpu) never links, loads or calls this code; only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark do, as the checker.
 *
 * Every function takes the same pcbgpu_query as the product's C ABI (include/pcbgpu.h).
 */
#ifndef PCB_ORACLE_H
#define PCB_ORACLE_H

#include "../include/pcbgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Resolved per-state cost factors of rectangle (rx0, ry0, rw, rh) of the window: cf (0 = may not be entered,
 * else [1, PCBGPU_CF_MAX]), vok (a via may land), tg (target); per cell vcf (the via cost, PCBGPU_INF: none). */
int orc_resolve(const pcbgpu_query* q, int rx0, int ry0, int rw, int rh, int32_t* cf, uint8_t* vok, uint8_t* tg,
                int32_t* vcf);
/* Dijkstra over the rectangle; d gets the canonical field, *best the rectangle state of the best target (-1),
 * *opt its cost (PCBGPU_INF: none). Returns 0 or -1. */
int orc_solve_rect(const pcbgpu_query* q, int rx0, int ry0, int rw, int rh, int32_t* d, int64_t* best, int32_t* opt);
/* The same by directional sweeps in the reference's pass order until nothing changes; returns the sweeps or -1. */
int orc_sweep_rect(const pcbgpu_query* q, int rx0, int ry0, int rw, int rh, int32_t* d, int64_t* best, int32_t* opt);
/* A whole solve with rectangle widening (the product's pcbgpu_solve). out->d is malloc'd: free with orc_out_free. */
int orc_solve(const pcbgpu_query* q, pcbgpu_out* out, size_t max_states);
void orc_out_free(pcbgpu_out* out);
/* The canonical backtrace, same contract as pcbgpu_backtrace. */
int64_t orc_backtrace(const pcbgpu_query* q, const pcbgpu_out* out, int32_t via_gapc, double via_gap2,
                      const uint32_t* old_vias, size_t n_old_vias, uint32_t* chain_states, uint8_t* chain_codes,
                      size_t chain_cap);
/* Source-to-target A* over the whole window with an octile heuristic to the targets' box (the CPU search the back
 * end finishes). Returns the cheapest target cost (PCBGPU_INF: none); *pops gets the heap pops. */
int32_t orc_astar(const pcbgpu_query* q, int64_t* pops);
/* Multi-threaded batch of A* searches (one query per task, nthreads workers). Returns 0. */
int orc_astar_batch(const pcbgpu_query* qs, int n, int nthreads, int32_t* opt, int64_t* pops);

#ifdef __cplusplus
}
#endif
#endif
