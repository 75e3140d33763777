"""Single-search kernel throughput on one MI355X (the 2000 x 1600 x 4 workload).

A step is the 64 seeded source / target searches of one synthetic board, each solved alone over the whole grid
(12.8M states) on a device-resident grid; the timed region starts with the maps already in HBM. Prints one JSON
line. --gpus N with N > 1 exits: this back end runs one search on one GPU.

Usage: python bench.py [--steps K] [--warmup W] [--pairs P] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
ALGO_BYTES = 12        # per state carried through one full relaxation: int32 cost read, int32 field read + write


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(board, pairs_seed, n_sample, threads):
    """The oracle's A* (the CPU search the back end finishes) on a bounded sample of the same searches."""
    from oracle import Oracle
    from pcbroute_amd import synth
    o = Oracle()
    probs = synth.pair_problems(board, pairs_seed, n_sample)
    t0 = time.perf_counter()
    opt, pops = o.astar_batch(probs, threads)
    dt = time.perf_counter() - t0
    return dict(value=round(n_sample / dt, 3), unit="searches/s", cores=threads, kind="port",
                sample=f"A* source-to-target on the first {n_sample} of the 64 searches, {threads} threads, "
                       f"{dt:.1f} s, {int(pops.sum())} pops"), opt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--board-seed", type=int, default=1)
    ap.add_argument("--pairs-seed", type=int, default=7)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--local-sweeps", type=int, default=0)
    ap.add_argument("--delta", type=int, default=-1, help="delta-stepping band over tiles (0: off)")
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=("single", "round"), default="single",
                    help="single: the searches one after another; round: all of a step's searches in one round")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the search back end solves each search on a single device; "
              "run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from pcbroute_amd import Backend, synth
    b = Backend()
    if not b.available():
        raise SystemExit("no HIP device")
    if a.tile_rows or a.local_sweeps:
        b.set_tile(a.tile_rows, a.local_sweeps)
    if a.delta >= 0:
        b.set_delta(a.delta)
    log("building the board")
    board = synth.kernel_board(a.board_seed)
    probs = synth.pair_problems(board, a.pairs_seed, a.pairs)
    nl, ny, nx = board[0].shape
    g = b.grid(probs[0])
    log(f"board {nx}x{ny}x{nl}, {len(probs)} searches; tile rows {b.get_tile(nl)}")

    def step():
        if a.mode == "round":
            return b.round([(g, p) for p in probs], field=False, rects=[None] * len(probs))
        res = []
        for p in probs:
            res.append(b.solve_rect_grid(g, p, field=False))
        return res

    for w in range(a.warmup):
        step()
        log(f"warmup {w + 1}/{a.warmup} done")
    b.stats(enable=True, reset=True)
    t0 = time.perf_counter()
    sweeps = []
    for k in range(a.steps):
        r = step()
        sweeps += [x.sweeps for x in r]
        log(f"step {k + 1}/{a.steps} done")
    wall = time.perf_counter() - t0
    st = b.stats(enable=False, reset=True)
    g.close()

    n_search = a.steps * len(probs)
    gcells = st["states"] / wall / 1e9                       # whole job: state relaxations per second
    kern_s = st["kernel_ms"] / 1e3
    achieved = st["states"] * ALGO_BYTES / kern_s / 1e9 if kern_s > 0 else 0.0
    line = {
        "metric": "grid-search Gcells/s (% HBM roof)",
        "value": round(gcells, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"single-search kernel throughput, plain: {nx}x{ny}x{nl} grid, "
                               f"{len(probs)} source/target pairs each solved alone over the whole grid",
                   "grid": [nx, ny, nl], "searches_per_step": len(probs), "obstacles": 0.2, "congested": 0.3,
                   "via_cost": synth.VIA, "move_weights": [synth.ORTH, synth.DIAG], "board_seed": a.board_seed,
                   "pairs_seed": a.pairs_seed, "mode": a.mode},
        "searches_per_s": round(n_search / wall, 3),
        "ms_per_search": round(wall / n_search * 1e3, 3),
        "passes_per_search": round(sum(sweeps) / len(sweeps), 2),
        "rect_state_passes_gcells_s": round(st["rect_state_passes"] / wall / 1e9, 3),
        "pct_hbm_roof": round(100 * achieved / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_pass", "launches": st["launches"],
                     "avg_launch_us": round(st["kernel_ms"] * 1e3 / max(1, st["launches"]), 3),
                     "algo_bytes_per_launch": round(st["states"] * ALGO_BYTES / max(1, st["launches"]))},
    }
    if not a.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cb, opt = cpu_baseline(board, a.pairs_seed, min(a.cpu_sample, a.pairs), threads)
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
