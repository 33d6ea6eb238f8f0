"""Writes tests/golden/golden.json: seeded config graphs (hash of the
generator output) with (total_scaled, diameter) computed by the reference's own
cpuASPLqueue template compiled from the reference tree (oracle/_ref/libref.so).

Run in the build container (where the reference tree exists):
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "graphgolf-cuda_amd"), os.path.join(REPO, "tests")]

import numpy as np  # noqa: E402

import graphgolf_mi355x as gg  # noqa: E402
import graphs  # noqa: E402
import oracle_lib as ol  # noqa: E402

CASES = [
    ("C1", "regular", dict(N=256, d=4, seed=s)) for s in range(1, 6)
] + [
    ("C2", "regular", dict(N=4096, d=12, seed=2)),
    ("C3", "orbit", dict(g=32, m=128, d=16, seed=3)),
    ("C4", "regular", dict(N=8192, d=16, seed=4)),
    ("C5", "regular", dict(N=16384, d=24, seed=5)),
]


def main():
    out = {"source": "oracle/_ref/libref.so = reference cppcode/cpuASPLqueue.cpp "
                     "(cpuASPLqueue<cap>::diameterASPL)", "cases": []}
    for name, kind, p in CASES:
        if kind == "regular":
            N, d = p["N"], p["d"]
            adj = gg.random_regular(N, d, p["seed"])
            h = hashlib.sha256(adj.tobytes()).hexdigest()
            diam, a = ol.ref_part(N, N, graphs.part_of_adj(adj))
            case = dict(name=name, kind=kind, N=N, d=d, seed=p["seed"], sha256=h)
        else:
            g, m, d = p["g"], p["m"], p["d"]
            N = g * m
            tab = gg.orbit_part(g, m, d, p["seed"])
            h = hashlib.sha256(tab.tobytes()).hexdigest()
            diam, a = ol.ref_part(N, m, tab)
            case = dict(name=name, kind=kind, N=N, g=g, m=m, d=d, seed=p["seed"], sha256=h)
        total = int(round(a * N * (N - 1)))
        assert float(total) / (N * (N - 1)) == a
        case.update(total_scaled=total, diameter=int(diam), aspl=a)
        out["cases"].append(case)
        print(name, case["total_scaled"], case["diameter"], flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
