"""Writes route_golden.json: the router (pcbroute_amd/csrc/router.cpp) over the CPU oracle's exact search
(oracle/libroute_oracle.so) on seeded synthetic boards — a 40x30 mm 2-layer board of 6 SOIC-16 and 20 0805 parts
(seed 1) and small 18x14 mm boards (seeds 1-3). Per board: the board hash, the route hash and the route's figures.
The HIP back end must reproduce them exactly. Run: python tests/golden/make_route_golden.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from pcbroute_amd import router as R  # noqa: E402

KEYS = ("nets", "connections", "opens", "violations", "conflict_nets", "iterations", "searches", "length_cells",
        "vias", "route_hash")


def main():
    lib = oracle.build_route_oracle()
    out = []
    for kind, seed in (("tiny", 1), ("tiny", 2), ("tiny", 3), ("c1", 1)):
        b = R.Board(R.spec(kind, seed=seed), lib_path=lib)
        r = R.Router(b, R.default_options(R.library(lib)), lib_path=lib)
        res = r.route()
        out.append({"kind": kind, "seed": seed, "board_hash": b.info.hash, **{k: res[k] for k in KEYS}})
        print(out[-1], flush=True)
    with open(os.path.join(HERE, "route_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
