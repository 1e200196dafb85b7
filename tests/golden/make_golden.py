"""Writes tests/golden/grids.json: process grids chosen by the reference's
own src/partition.c (compiled unmodified into oracle/_ref by oracle/Makefile)
for a fixed list of (p, m, n, k) problems, so that the grid search is pinned
on machines where the reference tree is absent.

Run from the repository root in a container that has the reference tree:
    make -C oracle && python tests/golden/make_golden.py
"""
import json
import os
import random
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def cases():
    rnd = random.Random(20221113)
    sizes = [1, 2, 3, 7, 10, 50, 100, 500, 1000, 2000, 2667, 4000, 6000, 8000, 8192,
             10007, 50000, 65536, 100000, 1200000, 2000000]
    out = [(24, 8000, 8000, 8000), (4, 2000, 2000, 2000), (1, 8192, 8192, 8192),
           (24, 10007, 10007, 10007), (2048, 50000, 50000, 50000), (2048, 6000, 6000, 1200000),
           (2048, 1200000, 6000, 6000), (2048, 100000, 100000, 5000)]
    for p in list(range(1, 33)) + [48, 64, 96, 128, 192, 256, 512, 1024, 2048, 3072]:
        for _ in range(12):
            out.append((p, rnd.choice(sizes), rnd.choice(sizes), rnd.choice(sizes)))
    return out


def main():
    r = O.ref_lib()
    if r is None:
        sys.exit("oracle/_ref/libca3dmm_ref.so not built (needs the reference tree)")
    cannon = [[p, m, n, k, *O.ref_grid_cannon(r, p, m, n, k)] for p, m, n, k in cases()]
    nk = []
    for p in [1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 100, 128, 256]:
        for n, k in [(512, 2000000), (500, 10000), (1000, 1000), (4096, 4096), (100, 1000000)]:
            nk.append([p, n, k, *O.ref_grid_nk(r, p, n, k)])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "grids.json")
    with open(path, "w") as f:
        json.dump({"source": "reference src/partition.c compiled unmodified (oracle/_ref)",
                   "cannon_fields": "p m n k -> mp np kp rp", "nk_fields": "p n k -> np kp rp",
                   "cannon": cannon, "nk": nk}, f, separators=(",", ":"))
    print(f"wrote {path}: {len(cannon)} cannon grids, {len(nk)} nk grids")


if __name__ == "__main__":
    main()
