"""Single-search kernel throughput on one MI355X (the 2000 x 1600 x 4 workload).

A step is the 64 seeded source / target searches of one synthetic board, each solved alone over the whole grid
(12.8M states, its own field) on a device-resident grid; the timed region starts with the maps already in HBM.
The searches of a step run as one round (one sequence of launches shared by all 64 searches, the batched path);
the same searches one after another (one solve at a time) are timed beside it. Prints one JSON line.

value: grid states solved per second (64 x 12.8M states per step / wall time of the round mode).
roofline: the sweep kernel k_pass, algorithmic bytes = 12 B per state per tile visit (int32 cost word read, int32
field read and write, once per visit however many local sweeps it runs) / summed kernel time from HIP events on the
launch stream, taken in as many steps run after the timed ones (the events launch the passes one by one; the timed
steps replay them from captured hipGraph chunks). Beside it: the HBM bytes rocprofv3's FETCH_SIZE and WRITE_SIZE counters measured per launch of the
same configuration (profiles/), their rate, and the figure that credits every local sweep with 12 B.
Tile rows, local sweeps and the delta band: the library's defaults for this workload (16-row tiles, up to 2 local
sweeps, automatic band of 20 straight steps), so the committed rocprof traffic profile matches the run; --tune auto runs the autotuner on
a sample drawn with another seed first. The batched round workload tunes by default.

--gpus N with N > 1 exits: the back end solves each search on a single device.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
ALGO_BYTES = 12        # per state-sweep (a state carried through the eight directional relaxations and the vias):
                       # int32 cost word read, int32 field read + write
DEFAULT_DELTA = 0      # delta-stepping band over tiles: 0 = the library's automatic band (20 straight steps)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(board, pairs_seed, n_sample, threads, n_total=64):
    """The oracle's A* (the CPU search the back end finishes) on a bounded sample of the same searches, and one
    full-field Dijkstra (what the back end computes) on one thread."""
    from oracle import Oracle
    from pcbroute_amd import synth
    o = Oracle()
    probs = synth.pair_problems(board, pairs_seed, n_sample)
    t0 = time.perf_counter()
    opt, pops = o.astar_batch(probs, threads)
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    o.solve_rect(probs[0])
    dfull = time.perf_counter() - t1
    return dict(value=round(n_sample / dt, 3), unit="searches/s", cores=threads, kind="port",
                sample=f"A* source-to-target (the CPU search the back end finishes) on the first {n_sample} of the "
                       f"{n_total} searches, {threads} threads, {dt:.1f} s, {int(pops.sum())} pops; one full-field "
                       f"Dijkstra of the same grid on 1 thread: {dfull:.2f} s",
                full_field_s=round(dfull, 3))


def run(b, g, probs, mode, steps, warmup):
    """The timed steps call the C ABI directly, as a C++ caller would, with each search's query built once before
    them (its inputs: the maps are on the device in the grid, the query holds the starts and the target bitmap)."""
    import ctypes as C
    from pcbroute_amd import NOFIELD
    from pcbroute_amd.abi import Out
    lib = b.lib
    qs = [p.query() for p in probs]

    def step():
        sweeps = []
        if mode == "round":
            r = lib.pcbgpu_round_begin()
            lib.pcbgpu_round_flags(r, NOFIELD)
            for p, q in zip(probs, qs):
                lib.pcbgpu_round_submit_rect(r, g.h, C.byref(q), 0, 0, p.window[2], p.window[3])
            rc = lib.pcbgpu_round_launch(r)
            for k in range(len(qs)):
                o = Out()
                rc = min(rc, lib.pcbgpu_round_result(r, k, C.byref(o)))
                sweeps.append(o.sweeps)
                lib.pcbgpu_out_release(C.byref(o))
            lib.pcbgpu_round_end(r)
        else:
            rc = 0
            for p, q in zip(probs, qs):
                o = Out()
                rc = min(rc, lib.pcbgpu_solve_rect_grid(g.h, C.byref(q), 0, 0, p.window[2], p.window[3], C.byref(o),
                                                        NOFIELD))
                sweeps.append(o.sweeps)
                lib.pcbgpu_out_release(C.byref(o))
        if rc < 0:
            raise RuntimeError(b.error())
        return sweeps

    for w in range(warmup):
        step()
        log(f"{mode}: warmup {w + 1}/{warmup} done")
    # the timed steps run the production path (a round replays its passes from captured hipGraph chunks); the kernel
    # accounting (HIP events around every pass, which launches passes one by one) runs in as many steps after them
    b.stats(enable=False, reset=True)
    t0 = time.perf_counter()
    sweeps = []
    for k in range(steps):
        sweeps += step()
        log(f"{mode}: step {k + 1}/{steps} done")
    wall = time.perf_counter() - t0
    b.stats(enable=True, reset=True)
    for k in range(steps):
        step()
    st = b.stats(enable=False, reset=True)
    log(f"{mode}: {steps} accounted steps done")
    return wall, st, sweeps


def round_with_fields(b, g, probs):
    """Wall time of one round whose canonical fields come back into the library's pinned output buffers (the
    pcbgpu_round_* calls directly: no copy into Python arrays); each search over its whole window as in the timed
    rounds."""
    import ctypes as C
    from pcbroute_amd.abi import Out
    lib = b.lib
    qs = [p.query() for p in probs]
    r = lib.pcbgpu_round_begin()
    for p, q in zip(probs, qs):
        lib.pcbgpu_round_submit_rect(r, g.h, C.byref(q), 0, 0, p.window[2], p.window[3])
    t0 = time.perf_counter()
    rc = lib.pcbgpu_round_launch(r)
    dt = time.perf_counter() - t0
    for k in range(len(qs)):
        o = Out()
        lib.pcbgpu_round_result(r, k, C.byref(o))
        lib.pcbgpu_out_release(C.byref(o))
    lib.pcbgpu_round_end(r)
    if rc != 0:
        raise RuntimeError(b.error())
    return dt


def tune(b, pairs, a):
    """The autotuner on a sample round (another seed than the timed searches); the chosen setting, or None."""
    if a.tune == "off":
        return None
    t0 = time.perf_counter()
    t = b.autotune(pairs, reps=2)
    log(f"autotune: {t} in {time.perf_counter() - t0:.1f} s")
    return {k: t[k] for k in ("tile_rows", "local_sweeps", "delta", "tried")}


def round_workload(b, a):
    """256 net searches of one speculative round (2-8 pins, windows log-uniform 64..2048 cells plus a margin) on a
    2400x1800x6 board with two net classes: one round, then the same searches one at a time in submission order;
    every result must agree bit for bit."""
    import numpy as np
    from pcbroute_amd import synth
    log("building the board (two net classes)")
    grids = synth.class_grids(a.board_seed)
    searches = synth.net_searches(grids, a.pairs_seed, 256)
    gs = [b.grid(synth_problem_of(g)) for g in grids]
    pairs = [(gs[k], p) for k, p in searches]
    states = sum(p.window[2] * p.window[3] * p.nl for _, p in searches)
    log(f"{len(pairs)} searches, {states / 1e6:.0f}M window states")
    tuned = tune(b, [(gs[k], p) for k, p in synth.net_searches(grids, a.pairs_seed + 1000, 64)], a)
    b.round(pairs, field=False)  # warmup
    b.stats(enable=True, reset=True)
    t0 = time.perf_counter()
    b.round(pairs, field=False)
    t_round = time.perf_counter() - t0
    st = b.stats(enable=False, reset=True)
    log("round done")
    t0 = time.perf_counter()
    for pr in pairs:  # one at a time: a round of one search is the lone solve
        b.round([pr], field=False)
    t_one = time.perf_counter() - t0
    log("one at a time done")
    # exactness (untimed): fields, rectangles, best targets and sweeps of the round against lone solves
    # (rectangle, best state and field; pass counts may differ by a few: the field is relaxed in place, in whatever
    # order workgroups reach the tiles)
    rr = b.round(pairs)
    same = True
    pass_diff = 0
    for (g, p), x in zip(pairs, rr):
        y = b.solve_grid(g, p)
        same = same and x.rect == y.rect and x.best == y.best and np.array_equal(x.d, y.d)
        pass_diff = max(pass_diff, abs(x.sweeps - y.sweeps))
    log("exactness check done")
    for g in gs:
        g.close()
    kern_s = st["kernel_ms"] / 1e3
    achieved = st["states"] * ALGO_BYTES / kern_s / 1e9 if kern_s > 0 else 0.0  # 12 B per state per tile visit
    print(json.dumps({
        "metric": "batched speculative round searches/s", "value": round(len(pairs) / t_round, 2),
        "unit": "searches/s", "n_gpus": 1, "higher_is_better": True, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "batched speculative round: 2400x1800x6, two net classes, 256 searches of 2-8 pins, "
                               "boxes log-uniform 64..2048 cells", "board_seed": a.board_seed,
                   "pairs_seed": a.pairs_seed, "window_states": states, "tuned": tuned},
        "round_s": round(t_round, 4), "one_at_a_time_s": round(t_one, 4), "speedup": round(t_one / t_round, 2),
        "bit_exact_with_one_at_a_time": bool(same), "max_pass_count_difference": pass_diff,
        "reached": int(sum(x.best >= 0 for x in rr)),
        "tries_total": int(sum(x.tries for x in rr)),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": "k_pass", "launches": st["launches"]},
    }), flush=True)


PRIOR_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix cores


def _ref_eval(ref, dsn, ses):
    import re
    import subprocess
    out = subprocess.run([ref, "-de", dsn, "--eval", ses], capture_output=True, text=True, timeout=600)
    m = re.search(r"eval: (\d+) nets, (\d+) connections, (\d+) open, (\d+) violations, ([\d.]+) mm of track, (\d+) vias",
                  out.stdout)
    if not m:
        return None
    k = ("nets", "connections", "open", "violations", "track_mm", "vias")
    return {n: (float(v) if n == "track_mm" else int(v)) for n, v in zip(k, m.groups())}


def c1_workload(b, a):
    """C1: the seeded 40x30 mm 2-layer board (50 nets) routed by the router on the GPU, one net at a time with full
    negotiation; the routes written as a session and judged by the reference's own evaluator; the reference's CPU
    router (oracle/_ref/pcbroute, built from its sources) routing the same board, written as DSN, timed beside it."""
    import subprocess
    import tempfile
    import numpy as np
    from pcbroute_amd import router as R
    board = R.Board(R.spec("c1", seed=a.board_seed))
    times, res, r = [], None, None
    for k in range(a.warmup + a.steps):
        if r is not None:
            r.close()
        r = R.Router(board)
        t0 = time.perf_counter()
        res = r.route()
        if k >= a.warmup:
            times.append(time.perf_counter() - t0)
        log(f"c1: route {k + 1}/{a.warmup + a.steps} done")
    ref = os.path.join(ROOT, "oracle", "_ref", "pcbroute")
    line = {
        "metric": "C1 board route time", "value": round(float(np.median(times)), 4), "unit": "s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(float(np.median(times)) * 1e3, 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "C1: seeded 40x30 mm 2-layer board, 6 SOIC-16 + 20 0805, 60 nets of 2-6 pins, "
                               "PathFinder, one net at a time", "board_seed": a.board_seed},
        "result": {k: res[k] for k in ("nets", "connections", "opens", "violations", "iterations", "searches",
                                       "length_cells", "vias", "route_hash")},
    }
    with tempfile.TemporaryDirectory() as d:
        dsn, ses = os.path.join(d, "c1.dsn"), os.path.join(d, "c1.ses")
        board.write_dsn(dsn)
        r.write_ses(ses)
        if os.path.exists(ref):
            line["reference_eval_of_our_session"] = _ref_eval(ref, dsn, ses)
            if not a.no_cpu:
                threads = min(16, os.cpu_count() or 1)
                out_ses = os.path.join(d, "ref.ses")
                t0 = time.perf_counter()
                p = subprocess.run([ref, "-de", dsn, "-do", out_ses, "-mt", str(threads)], capture_output=True,
                                   text=True, timeout=1200)
                t_ref = time.perf_counter() - t0
                q = _ref_eval(ref, dsn, out_ses) if p.returncode == 0 else None
                line["cpu_baseline"] = {"value": round(t_ref, 3), "unit": "s", "cores": threads, "kind": "reference",
                                        "sample": "the reference's CPU router (built from its sources) on the whole "
                                                  "board written as DSN, wall time of the process",
                                        "its_routes": q}
    r.close()
    print(json.dumps(line), flush=True)


def prior_workload(b, a):
    """The learned prior's training and inference on full 4-layer 60x50 mm boards: the 4x-downsampled grid is
    300 x 250 pixels, padded to 320 x 256; 22 input channels, 6 conv3x3 layers of 32 channels, 4 outputs. The
    training set stays in HBM; a step is one Adam step on a batch of 8 boards (forward, backward, update)."""
    import numpy as np
    from pcbroute_amd import prior as P
    h, w, cin, cout, batch, nds = 256, 320, 22, 4, 8, 16
    rng = np.random.default_rng(a.board_seed)
    x = rng.random((nds, h, w, cin), dtype=np.float32)
    t = (rng.random((nds, h, w, cout)) < 0.05).astype(np.float32)
    m = np.zeros((nds, h, w), np.float32)
    m[:, :250, :300] = 1
    net = P.PriorNet(cin, 32, 6, cout, seed=1)
    net.dataset(x, t, m)
    batches = [rng.integers(0, nds, batch).astype(np.int32) for _ in range(a.steps + a.warmup)]
    for k in range(a.warmup):
        net.train_resident(batches[k], loss=False)
    t0 = time.perf_counter()
    for k in range(a.steps):
        net.train_resident(batches[a.warmup + k], loss=False)
    wall = time.perf_counter() - t0
    log("prior: training timed")
    t0 = time.perf_counter()
    for k in range(a.steps):
        net.forward_resident(batches[k], logits=False)
    wall_inf = time.perf_counter() - t0
    net.stats(enable=True, reset=True)
    for k in range(a.steps):
        net.train_resident(batches[k], loss=False)
    st = net.stats(enable=False, reset=True)
    log("prior: kernel accounting done")
    px = float(batch) * h * w
    real = 250 * 300 / (h * w)
    step_flop = st["conv_flop"] / a.steps + st["wgrad_flop"] / a.steps
    # algorithmic HBM bytes (activations are 32 BF16 slots, 64 B per pixel): a forward conv reads its input and writes
    # its output (the last one writes cout FP32 logits); an input-gradient conv reads dy and the mask activation and
    # writes dx; a weight-gradient launch reads the layer's input and dy
    depth = 6
    conv_b = a.steps * px * ((depth - 1) * 128 + 64 + 4 * cout + (depth - 1) * 192)
    wgrad_b = a.steps * px * depth * 128
    kern = {"k_conv": (st["conv_flop"], st["conv_ms"], st["conv_launches"], conv_b),
            "k_wgrad": (st["wgrad_flop"], st["wgrad_ms"], st["wgrad_launches"], wgrad_b)}
    dom = max(kern, key=lambda k: kern[k][1])
    fl, ms, nlaunch, nbytes = kern[dom]
    achieved = fl / (ms / 1e3) / 1e12
    gbs = nbytes / (ms / 1e3) / 1e9
    line = {
        "metric": "prior training boards/s", "value": round(a.steps * batch / wall, 3), "unit": "boards/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "corridor prior: 6 conv3x3 layers 22->32x5->4 on 4-layer 60x50 mm boards "
                               "(300x250 prior pixels padded to 320x256), batch 8, Adam, training set resident in HBM",
                   "batch": batch, "pixels": [h, w], "channels": [cin, 32, cout]},
        "inference_boards_per_s": round(a.steps * batch / wall_inf, 3),
        "train_tflops": round(step_flop / (wall / a.steps) / 1e12, 2),
        "useful_fraction_of_pixels": round(real, 3),
        # both kernels sit below the ridge point (2.5 PFLOP/s / 8 TB/s = 312 FLOP/B): ~100-150 FLOP per algorithmic
        # byte, so HBM bounds them; the matrix-core fraction is given beside it
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbs / 8000.0, 4), "traffic": None, "kernel": dom,
                     "launches": nlaunch, "avg_launch_us": round(ms * 1e3 / max(1, nlaunch), 2),
                     "bytes_per_launch": round(nbytes / max(1, nlaunch)), "flop_per_launch": round(fl / max(1, nlaunch)),
                     "flop_per_byte": round(fl / nbytes, 1),
                     "mfma_tflops": round(achieved, 2), "mfma_peak": PRIOR_PEAK_TFLOPS,
                     "mfma_frac": round(achieved / PRIOR_PEAK_TFLOPS, 5)},
        "kernels": {k: {"tflops": round(v[0] / (v[1] / 1e3) / 1e12, 2), "mfma_frac": round(v[0] / (v[1] / 1e3) / 1e12
                                                                                       / PRIOR_PEAK_TFLOPS, 4),
                        "gb_s": round(v[3] / (v[1] / 1e3) / 1e9, 1), "hbm_frac": round(v[3] / (v[1] / 1e3) / 8e12, 4),
                        "ms": round(v[1], 3), "launches": v[2]}
                    for k, v in kern.items()},
    }
    if not a.no_cpu:
        from oracle import prior_ref as R
        params = net.unpack(net.get_params())
        t0 = time.perf_counter()
        R.loss_grad(params, x[:1], t[:1], m[:1], quant=False)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(1 / dt, 4), "unit": "boards/s", "cores": 1, "kind": "port",
                                "sample": f"one training step (forward + backward, float64 numpy) on 1 board of "
                                          f"320x256 pixels: {dt:.2f} s"}
    net.close()
    print(json.dumps(line), flush=True)


def c4board_workload(b, a):
    """C4 boards: held-out seeds routed plain and with the committed learned prior (pcbroute_amd.learn's
    configurations), every session judged by the reference's own evaluator and every board routed by the reference's
    CPU router (16 threads) beside them, where oracle/_ref/pcbroute is built. Prints one JSON line with the totals."""
    import subprocess
    import tempfile
    from pcbroute_amd import learn
    ref = os.path.join(ROOT, "oracle", "_ref", "pcbroute")
    net = learn.load_net()
    names = ["plain"] + [c for c in a.configs.split(",") if c and c != "plain"]
    with tempfile.TemporaryDirectory() as tmp:
        rows = learn.evaluate(net, [learn.EVAL_SEED0 + k for k in range(a.boards)],
                              configs={c: learn.CONFIGS[c] for c in names}, session_dir=tmp)
        tot = {c: {"route_s": 0.0, "opens": 0, "violations": 0, "judge_open": 0, "judge_violations": 0} for c in names}
        ref_tot = {"route_s": 0.0, "open": 0, "violations": 0}
        for row in rows:
            for c in names:
                x = row[c]
                tot[c]["route_s"] += x["route_s"]
                tot[c]["opens"] += x["opens"]
                tot[c]["violations"] += x["violations"]
                if os.path.exists(ref):
                    j = _ref_eval(ref, row["dsn"], x["ses"]) or {}
                    tot[c]["judge_open"] += j.get("open", 0)
                    tot[c]["judge_violations"] += j.get("violations", 0)
            if os.path.exists(ref) and a.reference_router:
                ses = os.path.join(tmp, f"{row['seed']}_reference.ses")
                t0 = time.perf_counter()
                p = subprocess.run([ref, "-de", row["dsn"], "-do", ses, "-mt", "16"], stdout=subprocess.DEVNULL,
                                   stderr=subprocess.DEVNULL, timeout=3600)
                ref_tot["route_s"] += time.perf_counter() - t0
                j = _ref_eval(ref, row["dsn"], ses) if p.returncode == 0 else None
                ref_tot["open"] += (j or {}).get("open", 0)
                ref_tot["violations"] += (j or {}).get("violations", 0)
            log(f"seed {row['seed']} judged")
    line = {"metric": "board route s, learned vs plain, same opens",
            "value": round(tot[names[-1]]["route_s"], 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
            "dtype": "int32", "data": "synthetic",
            "config": {"workload": f"C4: {a.boards} held-out 4-layer 60x50 mm boards (seeds {learn.EVAL_SEED0}..)",
                       "configs": names},
            "totals": {c: {k: round(v, 3) for k, v in t.items()} for c, t in tot.items()},
            "over_plain_time": {c: round(tot[c]["route_s"] / max(1e-9, tot["plain"]["route_s"]), 3) for c in names}}
    if a.reference_router and os.path.exists(ref):
        line["reference_cpu_router_16_threads"] = {k: round(v, 3) for k, v in ref_tot.items()}
    print(json.dumps(line), flush=True)


def c5board_workload(b, a):
    """C5: the 8-layer 160x120 mm board (two reserved inner planes, four BGA-256 whose nets escape in the 0.1 mm class,
    10 QFP-64, 30 SOIC-16, 400 0805, 1200 nets, three net classes) routed by the router on the GPU under a 240 s time
    limit, plain (one net at a time) and batched (each box-disjoint group of the order routed as one GPU round per
    step, commits validated in order); opens and violations by the router's own check and, where the reference's
    evaluator is built, by it. There is no learned prior for 8 layers: the network is trained for the 4-layer C4
    boards."""
    import subprocess
    import tempfile
    from pcbroute_amd import router as R
    board = R.Board(R.spec("c5", seed=a.board_seed))
    i = board.info
    log(f"c5 board {i.nx}x{i.ny}x{i.nl}: {i.n_nets} nets, {i.n_connections} connections, {i.n_parts} parts")
    ref = os.path.join(ROOT, "oracle", "_ref", "pcbroute")
    rows = {}
    with tempfile.TemporaryDirectory() as tmp:
        dsn = os.path.join(tmp, "c5.dsn")
        board.write_dsn(dsn)
        # no rip-ups while legalizing: under the time limit the rescue pass's pushes, each kept only if it opens fewer
        # connections, use the time after negotiation better (289 against 370 open, plain)
        for name, opts in (("plain", dict(wave=0, ripups=0)), ("batched", dict(wave=-64, ripups=0))):
            r = R.Router(board, R.default_options(time_limit_s=240.0, verbose=1, **opts))  # a line per round (stderr)
            t0 = time.perf_counter()
            res = r.route()
            wall = time.perf_counter() - t0
            opens, viol = r.eval()
            row = {"route_s": round(wall, 2), "opens": res["opens"], "violations": viol, "eval_opens": opens,
                   "iterations": res["iterations"], "searches": res["searches"], "gpu_rounds": res["rounds"],
                   "limited": res["limited"], "rescue_kept": res["rescue_kept"],
                   "opens_before_rescue": res["opens_before_rescue"], "vias": res["vias"],
                   "search_s": round(res["search_s"], 2)}
            if os.path.exists(ref):
                ses = os.path.join(tmp, name + ".ses")
                r.write_ses(ses)
                row["reference_judge"] = _ref_eval(ref, dsn, ses)
            rows[name] = row
            log(f"c5 {name}: {row}")
            r.close()
    print(json.dumps({
        "metric": "board route s, batched vs plain, same opens", "value": rows["batched"]["route_s"], "unit": "s",
        "n_gpus": 1, "higher_is_better": False, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "C5 board: 3200x2400x8 (two reserved planes), 4 BGA-256, 10 QFP-64, 30 SOIC-16, 400 "
                               "0805, three net classes, 240 s limit", "board_seed": a.board_seed,
                   "nets": i.n_nets, "connections": i.n_connections},
        "batched_over_plain_time": round(rows["batched"]["route_s"] / rows["plain"]["route_s"], 3),
        "plain": rows["plain"], "batched": rows["batched"]}), flush=True)


def synth_problem_of(grid):
    """A Problem that only carries a class's maps (to create its device-resident grid)."""
    from pcbroute_amd import Problem
    cst, st, vcst, vst = grid
    return Problem(cst, st, vcst, vst, starts=[], targets=[])


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
    ap.add_argument("--delta", type=int, default=DEFAULT_DELTA, help="delta-stepping band over tiles (0: automatic, -1: off)")
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=("both", "single", "round"), default="both")
    ap.add_argument("--workload", choices=("kernel", "round", "prior", "c1", "c4board", "c5", "c5board"),
                    default="kernel",
                    help="kernel: the 64 whole-grid searches (the metric); c5: the same on a 3200x2400x8 grid (the "
                         "C5 board's 61.4M states), 16 searches; round: 256 net searches of a speculative "
                         "round on a 2400x1800x6 board with two net classes, as one round and one at a time; "
                         "prior: the corridor CNN's training and inference on full 4-layer boards; c1: the router on "
                         "the C1 board, its session judged by the reference's evaluator, the reference's CPU router "
                         "timed beside it; c4board: held-out C4 boards, plain and learned (--configs), judged by the "
                         "reference's evaluator; c5board: the C5 board routed plain and batched under its time limit")
    ap.add_argument("--boards", type=int, default=8, help="c4board: held-out boards")
    ap.add_argument("--configs", default="plain,learned", help="c4board: pcbroute_amd.learn configurations")
    ap.add_argument("--reference-router", action="store_true", help="c4board: also route each board with the "
                    "reference's CPU router")
    ap.add_argument("--tune", choices=("auto", "off"), default=None,
                    help="auto: the autotuner picks tile rows, local sweeps and delta on a sample drawn with another "
                         "seed (the default for --workload round); off: library defaults or the explicit flags (the "
                         "default for the kernel workload)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 4 times (calibrates the HBM byte counters) and exit")
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        n = a.copy_probe * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(4):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": a.copy_probe, "copies": 4}), flush=True)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the search back end solves each search on a single device; "
              "run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from pcbroute_amd import Backend, synth
    b = Backend()
    if not b.available():
        raise SystemExit("no HIP device")
    if a.tune is None:
        a.tune = "auto" if a.workload == "round" else "off"
    if a.tile_rows or a.local_sweeps or a.delta != DEFAULT_DELTA:
        a.tune = "off"
    if a.tile_rows or a.local_sweeps:
        b.set_tile(a.tile_rows, a.local_sweeps)
    b.set_delta(a.delta)
    if a.workload == "round":
        return round_workload(b, a)
    if a.workload == "prior":
        return prior_workload(b, a)
    if a.workload == "c1":
        return c1_workload(b, a)
    if a.workload == "c5board":
        return c5board_workload(b, a)
    if a.workload == "c4board":
        return c4board_workload(b, a)
    log("building the board")
    dims = dict(nx=3200, ny=2400, nl=8) if a.workload == "c5" else {}
    if a.workload == "c5" and a.pairs == 64:
        a.pairs = 16
    board = synth.kernel_board(a.board_seed, **dims)
    probs = synth.pair_problems(board, a.pairs_seed, a.pairs)
    nl, ny, nx = board[0].shape
    g = b.grid(probs[0])
    tuned = tune(b, [(g, p) for p in synth.pair_problems(board, a.pairs_seed + 1000, 16)], a)
    rows, lsw = b.get_tile(nl)
    delta = tuned["delta"] if tuned else a.delta
    log(f"board {nx}x{ny}x{nl}, {len(probs)} searches; tile rows {rows}, local sweeps {lsw}, delta {delta}")
    modes = ("round", "single") if a.mode == "both" else (a.mode,)
    res = {m: run(b, g, probs, m, a.steps, a.warmup) for m in modes}
    g.close()

    head = modes[0]
    wall, st, sweeps = res[head]
    n_search = a.steps * len(probs)
    cells = nx * ny * nl
    kern_s = st["kernel_ms"] / 1e3
    # the algorithmic bytes: 12 per state per tile visit (cost read, field read, field write); a visit's local sweeps
    # work on the tile in registers and LDS, so crediting each of them with 12 B (the state-sweep figure) is reported
    # beside it only
    achieved = st["states"] * ALGO_BYTES / kern_s / 1e9 if kern_s > 0 else 0.0
    per_sweep = st["state_sweeps"] * ALGO_BYTES / kern_s / 1e9 if kern_s > 0 else 0.0
    avg_us = st["kernel_ms"] * 1e3 / max(1, st["launches"])
    line = {
        "metric": "grid-search Gcells/s (% HBM roof)",
        "value": round(n_search * cells / wall / 1e9, 3),
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
                   "pairs_seed": a.pairs_seed, "mode": head, "tile_rows": rows, "local_sweeps": lsw,
                   "delta": delta, "autotuned": tuned is not None},
        "searches_per_s": round(n_search / wall, 3),
        "ms_per_search": round(wall / n_search * 1e3, 3),
        "passes_per_search": round(sum(sweeps) / len(sweeps), 2),
        "tile_state_visits_per_search": round(st["states"] / n_search / cells, 2),
        "state_sweeps_per_search": round(st["state_sweeps"] / n_search / cells, 2),
        "gb_s_counting_each_local_sweep": round(per_sweep, 1),
        "processed_gstates_per_s": round(st["states"] / wall / 1e9, 3),
        "rect_state_passes_gcells_s": round(st["rect_state_passes"] / wall / 1e9, 3),
        "pct_hbm_roof": round(100 * achieved / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_pass", "launches": st["launches"],
                     "avg_launch_us": round(avg_us, 3),
                     "algo_bytes_per_launch": round(st["states"] * ALGO_BYTES / max(1, st["launches"]))},
    }
    # HBM traffic per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this configuration
    tpath = os.path.join(ROOT, "profiles", "k_pass_kernel_board_2000x1600x4_round.traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        want = {"tile_rows": rows, "local_sweeps": lsw, "delta": delta, "mode": head, "pairs": len(probs)}
        if t.get("config") == want:
            rl = line["roofline"]
            rl["traffic"] = t["traffic_bytes_per_launch"]
            rl["traffic_source"] = t["source"]
            # the counters' bytes at this run's launch time, and how far they exceed the algorithmic bytes
            rl["measured_gb_s"] = round(t["traffic_bytes_per_launch"] / (avg_us * 1e3), 1) if avg_us > 0 else None
            rl["measured_frac"] = round(rl["measured_gb_s"] / HBM_PEAK_GBS, 4) if avg_us > 0 else None
            rl["traffic_over_algorithmic"] = round(t["traffic_bytes_per_launch"] / max(1, rl["algo_bytes_per_launch"]), 3)
    # the same round with every search's canonical field copied back into pinned host memory (what a caller that
    # backtraces on the host pays): reported beside the value, never as it
    if a.workload == "kernel":
        g2 = b.grid(probs[0])
        t_fields = round_with_fields(b, g2, probs)
        line["with_fields"] = {"ms_per_search": round(t_fields / len(probs) * 1e3, 3),
                               "field_mb_per_search": round(cells * 4 / 1e6, 1)}
        g2.close()
    if "single" in res and head != "single":
        w1, s1, sw1 = res["single"]
        line["one_at_a_time"] = {"ms_per_search": round(w1 / n_search * 1e3, 3),
                                 "gcells_s": round(n_search * cells / w1 / 1e9, 3),
                                 "passes_per_search": round(sum(sw1) / len(sw1), 2)}
    if not a.no_cpu:
        # every host core this process may use, up to the 16 a one-GPU share of the box holds
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, min(16, avail))
        cb = cpu_baseline(board, a.pairs_seed, min(a.cpu_sample, a.pairs), threads, a.pairs)
        cb["nproc"] = os.cpu_count()
        cb["cores_available"] = avail
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
