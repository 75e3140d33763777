"""Run as its own process by tests/test_gpu_slots.py (the slot count takes effect only before the library's first
use): the slots and the state limit of pcbgpu_configure (src/gpu.cu:377-391, 550-561). Prints one JSON line."""
import json
import os
import sys
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

from pcbroute_amd import Backend, synth  # noqa: E402
from pcbroute_amd.abi import ERR_BUSY, ERR_FAIL, Out  # noqa: E402


def main():
    b = Backend()
    b.configure(int(sys.argv[1]), 0)  # before any solve: the slot count
    out = {"slots": int(sys.argv[1])}
    board = synth.kernel_board(2, nx=1200, ny=1000, nl=4)
    big = synth.pair_problems(board, 3, 4)
    small = synth.kernel_board(4, nx=96, ny=80, nl=2)
    sp = synth.pair_problems(small, 5, 1)[0]
    ref_small = b.solve(sp)
    # T threads at once, each its own searches, wait = 1: every result equals the one-thread result
    refs = [b.solve(p) for p in big]
    got = [None] * len(big)

    def work(k):
        got[k] = b.solve(big[k])
    th = [threading.Thread(target=work, args=(k,)) for k in range(len(big))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    out["concurrent_equal"] = all(g.rect == r.rect and g.best == r.best and np.array_equal(g.d, r.d)
                                  for g, r in zip(got, refs))
    # while every slot is held by long searches, a solve with wait = 0 returns busy (-3); with wait = 1 it waits
    import ctypes as C
    stop = threading.Event()

    def hold():
        for _ in range(40):
            for p in big:
                b.solve(p)
            if stop.is_set():
                break
    holders = [threading.Thread(target=hold) for _ in range(out["slots"])]
    for t in holders:
        t.start()
    time.sleep(0.05)  # the holders are inside their first searches
    busy = ok = 0
    q = sp.query()
    for _ in range(400):
        o = Out()
        rc = b.lib.pcbgpu_solve(C.byref(q), C.byref(o), 0)
        b.lib.pcbgpu_out_release(C.byref(o))
        if rc == ERR_BUSY:
            busy += 1
        elif rc >= 0:
            ok += 1
            # a poll that found a slot free took it; give the holder waiting for it time to take it back, or the
            # next poll wins the race again and the holders never run
            time.sleep(0.005)
        if busy >= 3:
            break
    waited = b.solve(sp, wait=True)
    stop.set()
    for t in holders:
        t.join()
    out["busy_seen"] = busy
    out["waited_equal"] = waited.rect == ref_small.rect and waited.best == ref_small.best and np.array_equal(waited.d, ref_small.d)
    # the state limit: a rectangle of more states than max_states is refused (-1), any time
    n_small = ref_small.rect[2] * ref_small.rect[3] * sp.nl  # the states of the rectangle it is solved on
    b.configure(0, n_small - 1)
    o = Out()
    rc = b.lib.pcbgpu_solve(C.byref(q), C.byref(o), 1)
    b.lib.pcbgpu_out_release(C.byref(o))
    out["over_limit_rc"] = rc
    out["over_limit_error"] = b.error()
    b.configure(0, n_small)
    out["at_limit_equal"] = b.solve(sp).best == ref_small.best
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
