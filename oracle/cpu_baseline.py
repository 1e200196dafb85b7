"""CPU baseline of the local DGEMM (TEST / BENCH INFRASTRUCTURE ONLY).

The reference's CPU path calls the host BLAS for its local DGEMM
(src/cannon.c:657-661, cblas_dgemm column-major NN, alpha = 1, beta = 0).
This module times the same call shape through numpy's BLAS (the image's
OpenBLAS) on a bounded sample, so bench.py can report it beside the GPU.
"""
import os
import time

import numpy as np


def blas_name():
    try:
        cfg = np.show_config(mode="dicts")
        return cfg["Build Dependencies"]["blas"]["name"]
    except Exception:  # pragma: no cover - numpy without config dicts
        return "numpy-blas"


def time_dgemm(n=4096, budget_s=12.0, seed=20221113):
    """Repeat an n^3 NN DGEMM (uniform[-1,1]) for about budget_s seconds.
    Returns (TFLOP/s, repetitions, seconds)."""
    rnd = np.random.default_rng(seed)
    A = np.asfortranarray(rnd.uniform(-1, 1, (n, n)))
    B = np.asfortranarray(rnd.uniform(-1, 1, (n, n)))
    Cm = np.empty((n, n), order="F")
    np.matmul(A, B, out=Cm)                     # warm-up (thread pool, pages)
    reps, t0 = 0, time.perf_counter()
    while True:
        np.matmul(A, B, out=Cm)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or reps >= 200:
            break
    return 2.0 * n ** 3 * reps / el / 1e12, reps, el


def threads():
    for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS"):
        v = os.environ.get(k)
        if v and v.isdigit():
            return int(v)
    return os.cpu_count() or 1
