"""Secondary measurements of bench.py (`python bench.py --suite NAME`).

Each suite prints JSON lines (one per measurement) on stdout:
  rank-share   one rank's share of the 4x2x3 grid on 8000^3: its two Cannon
               rounds (2000 x 4000 x 1334+1333), the kp = 3 partial-C fold
               over 2000 x 1334, and the mat_redist pack / unpack of the
               1D-column-partitioned 8000 x 8000 A for all 24 ranks (and rank
               0 alone), with achieved HBM GB/s on algorithmic bytes;
  ozaki-sweep  native FP64-MFMA DGEMM vs the two Ozaki INT8 forms at 8192^3
               and 65536 x 65536 x 1024 for phi in {0.5, 2, 4}, s in 3..11
               slices and 10..25 CRT moduli, error against double-double;
  ozaki-forms  native DGEMM, slices s = 8 and CRT with 15 moduli at n^3
               (phi = 0.5), three timed calls each: the kernel-trace workload;
  grouped      the simulated 4x2x3 grid on 10007^3: one launch per rank and
               round vs the persistent grouped launch (untuned / tuned / graph);
  cholqr       CholeskyQR of a 2e6 x 512 matrix at condition 1e2, 1e5, 1e8, 1e12 (or --cond): each variant;
  pmc-calib    a plain 1 GiB copy, then one of each hot kernel, for counter runs;
  dgemm-group  hip_dgemm at n^3 per tile-row group size (CA3DMM_DGEMM_GROUP);
  yardstick    rocBLAS dgemm and INT8 gemm_ex at the same size (never on the path).
"""
import json
import os
import sys
import time

import numpy as np

from ca3dmm_amd import device as hip
from ca3dmm_amd import engine as E

HBM_GBS = 8000.0
FP64_TF = 78.6


def emit(d):
    print(json.dumps(d), flush=True)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def phi_matrix(rnd, shape, phi):
    return _bench().phi_matrix(rnd, shape, phi)


def dd_sample(transA, transB, A, B, ii, jj):
    """(ref, ab) of the sampled entries in double-double (bench.dd_sample)."""
    return _bench().dd_sample(transA, transB, A, B, ii, jj)


def _bench():
    import bench
    return bench


def col1d(nrow, ncol, p):
    """example_AB's 1D column layout (srow, nrow, scol, ncol) per rank: the
    first ncol % p ranks own one column more (calc_block_spos_size,
    src/utils.c:26-49)."""
    q, rem = divmod(ncol, p)
    return [(0, nrow, r * q + min(r, rem), q + (1 if r < rem else 0)) for r in range(p)]


def grid_redist(m, n, k, p, proc_grid=None):
    """The redistribution plans of A and B that the library builds for the
    p-rank grid on m x n x k with the 1D column layout, read back from its
    engines: {name: {"src", "req", "plans"}} with plans[r] = (send list,
    send count, receive list, receive count), list rows (srow, scol, nrow,
    ncol, peer, displ) in the engines' row-major view (rows = columns)."""
    from ca3dmm_amd import engine as E
    sim = E.SimGrid(p, m, n, k, 0, 0, col1d(m, k, p), col1d(k, n, p), col1d(m, n, p), proc_grid=proc_grid,
                    allocate=False)
    out = {}
    try:
        for name in ("A", "B"):
            src, req, plans = [], [], []
            for r in range(p):
                e = sim.engine(r)
                rd = (e.redist_A if name == "A" else e.redist_B).contents
                src.append((rd.src_srow, rd.src_scol, rd.src_nrow, rd.src_ncol))
                req.append((rd.req_srow, rd.req_scol, rd.req_nrow, rd.req_ncol))
                plans.append((rd.send_list().tolist(), rd.send_cnt, rd.recv_list().tolist(), rd.recv_cnt))
            out[name] = {"src": src, "req": req, "plans": plans}
    finally:
        sim.free()
    return out


def check_blocks(blocks, bufs, dt=8):
    """Host-side bounds check of row-major block copies against the
    allocations they read and write: every block must lie inside one buffer."""
    spans = sorted((b.ptr, b.ptr + b.nbytes) for b in bufs)
    def inside(addr, nr, nc, ld):
        end = addr + dt * ((nr - 1) * ld + nc)
        return any(lo <= addr and end <= hi for lo, hi in spans)
    for s, lds, d, ldd, nr, nc in blocks:
        if nr <= 0 or nc <= 0:
            continue
        if not (nc <= lds and nc <= ldd and inside(s, nr, nc, lds) and inside(d, nr, nc, ldd)):
            raise ValueError(f"block copy out of bounds: {(s, lds, d, ldd, nr, nc)}")


def timed(f, reps=5, warm=1):
    for _ in range(warm):
        f()
    hip.sync()
    t = hip.Timer()
    t.start()
    for _ in range(reps):
        f()
    return t.stop() / reps


# ---------------------------------------------------------------------------
def rank_share(args):
    hip.set_sync(False)
    rnd = np.random.default_rng(1)
    # two Cannon rounds of rank 0: C(2000 x 4000) = A0 B0 + A1 B1, k 1334 + 1333
    m, n, k0, k1 = 2000, 4000, 1334, 1333
    A0, A1 = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (m, k0))), hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (m, k1)))
    B0, B1 = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (k0, n))), hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (k1, n)))
    Cm = hip.DeviceArray.zeros((m, n))

    def rounds():
        hip.dgemm(0, 0, m, n, k0, 1.0, A0, B0, 0.0, Cm)
        hip.dgemm(0, 0, m, n, k1, 1.0, A1, B1, 1.0, Cm)
    ms = timed(rounds, reps=10)
    fl = 2.0 * m * n * (k0 + k1)
    emit({"suite": "rank-share", "what": "rank 0 Cannon rounds 2000x4000x(1334+1333)", "ms": round(ms, 4),
          "tflops": round(fl / ms / 1e9, 2), "frac_fp64_peak": round(fl / ms / 1e9 / FP64_TF, 4)})
    # kp = 3 fold of the 2000 x 1334 slice
    P = [hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (2000, 1334))) for _ in range(3)]
    out = hip.DeviceArray.zeros((2000, 1334))
    ms = timed(lambda: hip.fold_sum(P, 2000, 1334, out), reps=50)
    by = 4 * 8 * 2000 * 1334
    emit({"suite": "rank-share", "what": "kp=3 fold 2000x1334", "ms": round(ms, 5), "bytes": by,
          "gbs": round(by / ms / 1e6, 1), "frac_hbm": round(by / ms / 1e6 / HBM_GBS, 4)})
    # the same fold on the whole 8000 x 8000 C (all ranks' slices), one launch
    P8 = [hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (8000, 8000))) for _ in range(3)]
    out8 = hip.DeviceArray.zeros((8000, 8000))
    ms = timed(lambda: hip.fold_sum(P8, 8000, 8000, out8), reps=10)
    by = 4 * 8 * 8000 * 8000
    emit({"suite": "rank-share", "what": "kp=3 fold 8000x8000", "ms": round(ms, 4), "bytes": by,
          "gbs": round(by / ms / 1e6, 1), "frac_hbm": round(by / ms / 1e6 / HBM_GBS, 4)})
    del P8, out8
    # pack / unpack of A for all 24 ranks (example_AB 1D column layout -> cuboid layout)
    plans = grid_redist(8000, 8000, 8000, 24)
    red = plans["A"]
    Afull = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (8000, 8000)))
    sends = [hip.DeviceArray((max(1, red["plans"][r][1]),)) for r in range(24)]
    recvs = [hip.DeviceArray((max(1, red["plans"][r][3]),)) for r in range(24)]
    dsts = [hip.DeviceArray((max(1, red["req"][r][2] * red["req"][r][3]),)) for r in range(24)]
    pack, unpack, pack0 = [], [], []
    for r in range(24):
        s = [int(v) for v in red["src"][r]]
        sl, sc, rl, rc = red["plans"][r]
        sl = [[int(v) for v in e] for e in sl]
        rl = [[int(v) for v in e] for e in rl]
        base = Afull.at(s[0] * 8000)        # rank r's column block (row-major rows = columns)
        for e in sl:
            d = (base + 8 * ((e[0] - s[0]) * 8000 + (e[1] - s[1])), 8000, sends[r].at(e[5]), e[3], e[2], e[3])
            pack.append(d)
            if r == 0:
                pack0.append(d)
        q = [int(v) for v in red["req"][r]]
        for e in rl:
            unpack.append((recvs[r].at(e[5]), e[3], dsts[r].at((e[0] - q[0]) * q[3] + (e[1] - q[1])), q[3], e[2], e[3]))
    check_blocks(pack, [Afull] + sends)
    check_blocks(unpack, recvs + dsts)
    # the one-pass redistribution of A and B for all 24 ranks (what the
    # simulated grid runs): each block from the sender's column block straight
    # into the receiver's redistribution buffer, both matrices in one call
    Bfull = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (8000, 8000)))
    direct, direct0 = [], []
    bufs = [Afull, Bfull]
    for name, full in (("A", Afull), ("B", Bfull)):
        red = plans[name]
        dst = [hip.DeviceArray((max(1, red["req"][r][2] * red["req"][r][3]),)) for r in range(24)]
        bufs += dst
        for r in range(24):
            q = [int(v) for v in red["req"][r]]
            for e in red["plans"][r][2]:
                e = [int(v) for v in e]
                s0 = [int(v) for v in red["src"][e[4]]]
                src = full.at(s0[0] * 8000) + 8 * ((e[0] - s0[0]) * 8000 + (e[1] - s0[1]))
                d = (src, 8000, dst[r].at((e[0] - q[0]) * q[3] + (e[1] - q[1])), q[3], e[2], e[3])
                direct.append(d)
                if r == 0:
                    direct0.append(d)
    check_blocks(direct, bufs)
    for name, lst in (("pack A, 24 ranks", pack), ("unpack A, 24 ranks", unpack), ("pack A, rank 0", pack0),
                      ("one-pass redistribution of A and B, 24 ranks", direct),
                      ("one-pass redistribution of A and B, rank 0", direct0)):
        elems = sum(d[4] * d[5] for d in lst)
        ms = timed(lambda: hip.copy_blocks(8, lst), reps=20)
        by = 16 * elems
        emit({"suite": "rank-share", "what": name, "blocks": len(lst), "elements": elems, "ms": round(ms, 5),
              "bytes": by, "gbs": round(by / ms / 1e6, 1), "frac_hbm": round(by / ms / 1e6 / HBM_GBS, 4)})
    # the copy ceiling at rank 0's sizes: one contiguous run of the same
    # element count through the same copy kernel (launch and ramp included)
    for name, lst in (("pack A, rank 0", pack0), ("one-pass redistribution of A and B, rank 0", direct0)):
        elems = sum(d[4] * d[5] for d in lst)
        src, dst = hip.DeviceArray((elems,)), hip.DeviceArray((elems,))
        one = [(src.ptr, elems, dst.ptr, elems, 1, elems)]
        ms_copy = timed(lambda: hip.copy_blocks(8, one), reps=20)
        ms_blk = timed(lambda: hip.copy_blocks(8, lst), reps=20)
        by = 16 * elems
        emit({"suite": "rank-share", "what": f"{name} against a plain copy of its {elems} elements",
              "ms_blocks": round(ms_blk, 5), "ms_plain_copy": round(ms_copy, 5),
              "gbs_plain_copy": round(by / ms_copy / 1e6, 1), "blocks_over_plain": round(ms_copy / ms_blk, 3)})
        del src, dst
    del bufs, direct, direct0
    # the whole 8000^3 grid in one persistent grouped launch (all 24 ranks'
    # Cannon rounds and the kp = 3 fold), autotuned, then the full exec
    # (redistributions included) replayed from its hipGraph
    from ca3dmm_amd import engine as E
    A8 = rnd.uniform(-1, 1, (8000, 8000))
    B8 = rnd.uniform(-1, 1, (8000, 8000))
    dA8, dB8 = hip.DeviceArray.from_numpy(A8), hip.DeviceArray.from_numpy(B8)
    dC8 = hip.DeviceArray.zeros((8000, 8000))
    rects = col1d(8000, 8000, 24)
    sim = E.SimGrid(24, 8000, 8000, 8000, 0, 0, rects, rects, rects, proc_grid=[4, 2, 3])
    Ap = [dA8.at(r[2] * 8000) for r in rects]
    Bp = [dB8.at(r[2] * 8000) for r in rects]
    Cp = [dC8.at(r[2] * 8000) for r in rects]
    run = lambda: sim.exec(Ap, [8000] * 24, Bp, [8000] * 24, Cp, [8000] * 24)
    fl = 2.0 * 8000 ** 3
    ms = timed(run, reps=5)
    emit({"suite": "rank-share", "what": "8000^3 on 4x2x3, per-rank mode (all ranks' rounds in one batched launch), whole exec",
          "ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 2)})
    sim.set_mode(grouped=True, graph=False)
    g, wg, tile = sim.autotune(3)
    ms = timed(run, reps=5)
    tname = {0: "128x128", 1: "256x128", 2: "128x256", 3: "batched 128x128 direct-to-LDS"}[tile]
    emit({"suite": "rank-share", "what": f"8000^3 on 4x2x3, grouped (tile {tname}, rows={g}, {wg} WG/CU), whole exec",
          "ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 2)})
    sim.set_mode(grouped=True, graph=True)
    ms = timed(run, reps=5)
    emit({"suite": "rank-share", "what": "8000^3 on 4x2x3, grouped, whole exec replayed from one hipGraph",
          "ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 2)})
    sim.free()
    hip.set_sync(True)


# ---------------------------------------------------------------------------
def ozaki_sweep(args):
    """Error on 4096 sampled entries against double-double, per (shape, phi,
    form, count): the slice form for s = 3..11 and the CRT form for 10..25
    moduli; `no_worse_than_native` needs both error measures at or below the
    native product's."""
    rnd = np.random.default_rng(7)
    shapes = [(args.n, args.n, args.n), (65536, 65536, 1024)]
    forms = [("ozaki-int8", "s", range(3, 12), hip.ozaki_dgemm, lambda s: s * (s + 1) // 2),
             ("ozaki-int8-crt", "moduli", range(10, 26), hip.ozaki2_dgemm, lambda c: c)]
    for m, n, k in shapes:
        flops = 2.0 * m * n * k
        for phi in (0.5, 2.0, 4.0):
            A, B = phi_matrix(rnd, (m, k), phi), phi_matrix(rnd, (k, n), phi)
            dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
            dC = hip.DeviceArray.zeros((m, n))
            ii, jj = rnd.integers(0, m, 4096), rnd.integers(0, n, 4096)
            ref, ab = dd_sample(0, 0, A, B, ii, jj)
            del A, B
            hip.set_sync(False)
            nat_ms = ms = timed(lambda: hip.dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC), reps=2)
            got = dC.sample(ii, jj)
            nat = float(np.max(np.abs(got - ref) / ab))
            nat_rel = float(np.max(np.abs(got - ref) / np.abs(ref)))
            base = {"suite": "ozaki-sweep", "m": m, "n": n, "k": k, "phi": phi, "samples": 4096}
            emit(dict(base, method="dgemm-mfma", ms=round(ms, 3), tflops=round(flops / ms / 1e9, 2),
                      max_err_over_absAabsB=nat, max_rel_err=nat_rel))
            for name, key, counts, fn, products in forms:
                for c in counts:
                    ms = timed(lambda: fn(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC, c), reps=2)
                    got = dC.sample(ii, jj)
                    err = float(np.max(np.abs(got - ref) / ab))
                    rel = float(np.max(np.abs(got - ref) / np.abs(ref)))
                    emit(dict(base, method=name, **{key: c}, ms=round(ms, 3),
                              effective_tflops=round(flops / ms / 1e9, 2),
                              int8_tops=round(products(c) * flops / ms / 1e9, 1),
                              max_err_over_absAabsB=err, max_rel_err=rel,
                              no_worse_than_native=(err <= nat and rel <= nat_rel),
                              speedup_vs_native=round(nat_ms / ms, 3)))
            hip.set_sync(True)
            del dA, dB, dC


def ozaki_forms(args):
    """One shape, the three FP64-accurate products side by side (for kernel
    traces): native FP64-MFMA DGEMM, the slice form at s = 8 and the CRT
    form at 15 moduli, n x n x k (k = --k, default n) with phi = 0.5 inputs;
    one warm-up and three timed calls each, error on 1024 sampled entries
    against double-double."""
    n = args.n
    k = args.k or n
    rnd = np.random.default_rng(1)
    A, B = phi_matrix(rnd, (n, k), 0.5), phi_matrix(rnd, (k, n), 0.5)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.zeros((n, n))
    ii, jj = rnd.integers(0, n, 1024), rnd.integers(0, n, 1024)
    ref, ab = dd_sample(0, 0, A, B, ii, jj)
    del A, B
    flops = 2.0 * n * n * k
    hip.set_sync(False)
    for name, f, prods in (("dgemm-mfma", lambda: hip.dgemm(0, 0, n, n, k, 1.0, dA, dB, 0.0, dC), None),
                           ("ozaki-int8 s=8", lambda: hip.ozaki_dgemm(0, 0, n, n, k, 1.0, dA, dB, 0.0, dC, 8), 36),
                           ("ozaki-int8-crt moduli=15",
                            lambda: hip.ozaki2_dgemm(0, 0, n, n, k, 1.0, dA, dB, 0.0, dC, 15), 15)):
        ms = timed(f, reps=3)
        got = dC.sample(ii, jj)
        e = np.abs(got - ref)
        rec = {"suite": "ozaki-forms", "n": n, "k": k, "phi": 0.5, "method": name, "ms": round(ms, 3),
               "effective_tflops": round(flops / ms / 1e9, 2), "max_err_over_absAabsB": float(np.max(e / ab)),
               "max_rel_err": float(np.max(e / np.abs(ref)))}
        if prods:
            rec["int8_tops"] = round(prods * flops / ms / 1e9, 1)
        emit(rec)
    hip.set_sync(True)


# ---------------------------------------------------------------------------
def grouped(args):
    """The simulated 4x2x3 grid on 10007^3, whole sim_exec per call (both
    redistributions, transposes, allgather, products, k-layer fold):
    FP64 form per launch / grouped untuned / tuned / hipGraph, then the
    Ozaki CRT form per rank / grouped / hipGraph, with the error of every
    form on 512 sampled entries against double-double."""
    m = n = k = 10007
    p = 24
    rnd = np.random.default_rng(3)
    lay_A, lay_B, lay_C = col1d(m, k, p), col1d(k, n, p), col1d(m, n, p)
    A = rnd.uniform(-1, 1, (m, k))
    B = rnd.uniform(-1, 1, (k, n))
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    ii, jj = rnd.integers(0, m, 512), rnd.integers(0, n, 512)
    ref, ab = dd_sample(0, 0, A, B, ii, jj)
    del A, B

    def errs():
        got = dC.sample(ii, jj)
        e = np.abs(got - ref)
        return {"max_err_over_absAabsB": float(np.max(e / ab)), "max_rel_err": float(np.max(e / np.abs(ref)))}
    dC = hip.DeviceArray.zeros((m, n))
    Ap = [dA.at(r[2] * m) for r in lay_A]
    Bp = [dB.at(r[2] * k) for r in lay_B]
    Cp = [dC.at(r[2] * m) for r in lay_C]
    sim = E.SimGrid(p, m, n, k, 0, 0, lay_A, lay_B, lay_C, proc_grid=[4, 2, 3])
    flops = 2.0 * m * n * k
    run = lambda: sim.exec(Ap, [m] * p, Bp, [k] * p, Cp, [m] * p)
    hip.set_sync(False)
    ms = timed(run, reps=3)
    ref_c = dC.numpy()
    emit(dict({"suite": "grouped", "mode": "per-rank: all ranks' rounds in one batched launch", "ms": round(ms, 3),
               "tflops": round(flops / ms / 1e9, 2)}, **errs()))
    sim.set_mode(grouped=True, graph=False)
    sim.set_tuning(1, 1)
    ms = timed(run, reps=3)
    d = float(np.max(np.abs(dC.numpy() - ref_c)))
    emit({"suite": "grouped", "mode": "grouped, untuned (rows=1, 1 WG/CU)", "ms": round(ms, 3),
          "tflops": round(flops / ms / 1e9, 2), "max_abs_diff_vs_per_launch": d})
    g, wg, tile = sim.autotune(2)
    ms = timed(run, reps=3)
    tname = {0: "128x128", 1: "256x128", 2: "128x256", 3: "batched 128x128 direct-to-LDS"}[tile]
    emit({"suite": "grouped", "mode": f"grouped, tuned (tile {tname}, rows={g}, {wg} WG/CU)", "ms": round(ms, 3),
          "tflops": round(flops / ms / 1e9, 2)})
    sim.set_mode(grouped=True, graph=True)
    ms = timed(run, reps=3)
    emit(dict({"suite": "grouped", "mode": "grouped, tuned, hipGraph replay", "ms": round(ms, 3),
               "tflops": round(flops / ms / 1e9, 2),
               "max_abs_diff_vs_per_launch": float(np.max(np.abs(dC.numpy() - ref_c)))}, **errs()))
    for nmod in (15, 14, 16):
        sim.set_ozaki(nmod)

        def rec(mode):
            ms = timed(run, reps=3)
            emit(dict({"suite": "grouped", "form": f"ozaki-int8-crt moduli={nmod}", "mode": mode,
                       "ms": round(ms, 3), "effective_tflops": round(flops / ms / 1e9, 2)}, **errs()))
        if nmod == 15:
            sim.set_tuning_ozaki(256, 0)
            sim.set_mode(grouped=False, graph=False)
            rec("per-rank product launches, tile 256x256")
            sim.set_mode(grouped=True, graph=False)
            rec("grouped, fixed tile 256x256, single queue")
        sim.set_mode(grouped=True, graph=False)
        t, q = sim.autotune_ozaki(2)
        tuned = f"tile 256x{t}, {'per-XCD' if q else 'single'} queue"
        if nmod == 15:
            rec(f"grouped, tuned ({tuned})")
        sim.set_mode(grouped=True, graph=True)
        rec(f"grouped, tuned ({tuned}), hipGraph replay")
    sim.set_ozaki(0)
    hip.set_sync(True)
    sim.free()


# ---------------------------------------------------------------------------
def cholqr(args):
    """CholeskyQR of 2e6 x 512 at four condition numbers (or --cond) through
    hip_cholesky_qr: CholeskyQR2 with the Gram through FP64 MFMA and through
    Ozaki INT8 (s = 7, 8), shifted CholeskyQR3, and the automatic variant."""
    from ca3dmm_amd import cholqr as CQ
    m, n = 2000000, 512
    dU = None
    for cond in ((args.cond,) if args.cond else (1e2, 1e5, 1e8, 1e12)):
        t0 = time.perf_counter()
        dA, dU = CQ.test_matrix(m, n, cond, seed=2022, dU=dU)
        log(f"cholqr: matrix kappa={cond:g} built in {time.perf_counter() - t0:.1f} s")
        dG = hip.DeviceArray.zeros((n, n))
        cases = [(hip.CHOLQR2, None), (hip.CHOLQR2, 7), (hip.CHOLQR2, 8), (hip.SCHOLQR3, None),
                 (hip.CHOLQR_AUTO, None)]
        for variant, gs in cases:
            hip.set_sync(False)
            if gs:
                ms_gram = timed(lambda: hip.ozaki_gram(n, m, dA, dG, gs), reps=3)
            else:
                ms_gram = timed(lambda: hip.gram(n, m, dA, dG), reps=3)
            hip.set_sync(True)
            name = {hip.CHOLQR2: "cholqr2", hip.SCHOLQR3: "shifted cholqr3", hip.CHOLQR_AUTO: "auto"}[variant]
            rec = {"suite": "cholqr", "m": m, "n": n, "cond": cond, "variant": name,
                   "gram": f"ozaki-int8 s={gs}" if gs else "fp64-mfma", "gram_ms": round(ms_gram, 3),
                   "gram_tflops": round(n * (n + 1) * m / ms_gram / 1e9, 2)}
            ms, dQ, dR, passes, info = _bench().time_cholesky_qr(hip, dA, m, n, variant, gram_slices=gs or 0, reps=1)
            if info:
                rec.update({"breakdown": f"Cholesky breakdown in pass {passes + 1}", "info": info, "pass": passes + 1})
            else:
                rec["ms"] = round(ms, 2)
                rec["passes"] = passes
                orth, res = CQ.qr_errors(dA, dQ, dR, m, n)
                rec.update({"orth_QtQ_minus_I_F": orth, "residual_rel_F": res})
            del dQ, dR
            emit(rec)
        del dA, dG


# ---------------------------------------------------------------------------
def pmc_calib(args):
    """Counter-run program: a plain copy of 1 GiB (the scale of FETCH_SIZE /
    WRITE_SIZE), then one launch of each hot kernel at its bench size: the
    four DGEMM trans cases at n^3 (direct-to-LDS stages, the default), the
    kp = 3 fold over 2000 x 1334, the pack of A for the 24 ranks of the 4x2x3
    grid on 8000^3, a 7-slice Ozaki product n^3 and a 15-modulus CRT-form
    product n^3; then the NN DGEMM register-staged throughout, the fold over
    8000 x 8000, one 8000^3 grid exec per-rank and one grouped, and the Gram
    of a 2e6 x 512 operand."""
    nel = (1 << 30) // 8
    src = hip.DeviceArray.zeros((nel,))
    dst = hip.DeviceArray.zeros((nel,))
    hip.copy_matrix_block(8, 1, nel, src.ptr, nel, dst.ptr, nel)
    hip.sync()
    del src, dst
    n = args.n
    rnd = np.random.default_rng(5)
    dA = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n, n)))
    dB = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n, n)))
    dC = hip.DeviceArray.zeros((n, n))
    for ta, tb in ((0, 0), (0, 1), (1, 0), (1, 1)):
        hip.dgemm(ta, tb, n, n, n, 1.0, dA, dB, 0.0, dC)
    P = [hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (2000, 1334))) for _ in range(3)]
    out = hip.DeviceArray.zeros((2000, 1334))
    hip.fold_sum(P, 2000, 1334, out)
    red = grid_redist(8000, 8000, 8000, 24)["A"]
    rects = col1d(8000, 8000, 24)
    Afull = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (8000, 8000)))
    sends = [hip.DeviceArray((max(1, red["plans"][r][1]),)) for r in range(24)]
    pack = []
    for r in range(24):
        s0 = [int(v) for v in red["src"][r]]
        base = Afull.at(s0[0] * 8000)
        for e in ([int(v) for v in e] for e in red["plans"][r][0]):
            pack.append((base + 8 * ((e[0] - s0[0]) * 8000 + (e[1] - s0[1])), 8000, sends[r].at(e[5]),
                         e[3], e[2], e[3]))
    check_blocks(pack, [Afull] + sends)
    hip.copy_blocks(8, pack)
    hip.ozaki_dgemm(0, 0, n, n, n, 1.0, dA, dB, 0.0, dC, 7)
    hip.ozaki2_dgemm(0, 0, n, n, n, 1.0, dA, dB, 0.0, dC, 15)
    hip.sync()
    # the same DGEMM with register-staged stages throughout (NN)
    os.environ["CA3DMM_DGEMM_LOADS"] = "reg"
    hip.dgemm(0, 0, n, n, n, 1.0, dA, dB, 0.0, dC)
    hip.sync()
    del os.environ["CA3DMM_DGEMM_LOADS"]
    del dA, dB, dC, P, out, sends
    # the kp = 3 fold over the whole 8000 x 8000 C
    P8 = [hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (8000, 8000))) for _ in range(3)]
    out8 = hip.DeviceArray.zeros((8000, 8000))
    hip.fold_sum(P8, 8000, 8000, out8)
    hip.sync()
    del P8, out8
    # the 8000^3 grid on 4x2x3: one exec per-rank (one-pass redistributions of
    # A and B in one launch, 48 DGEMM rounds, 24 folds, C redistribution),
    # then one exec through the grouped launch
    from ca3dmm_amd import engine as E
    dA8 = Afull
    dB8 = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (8000, 8000)))
    dC8 = hip.DeviceArray.zeros((8000, 8000))
    sim = E.SimGrid(24, 8000, 8000, 8000, 0, 0, rects, rects, rects, proc_grid=[4, 2, 3])
    Ap = [dA8.at(r[2] * 8000) for r in rects]
    Bp = [dB8.at(r[2] * 8000) for r in rects]
    Cp = [dC8.at(r[2] * 8000) for r in rects]
    sim.exec(Ap, [8000] * 24, Bp, [8000] * 24, Cp, [8000] * 24)
    sim.set_mode(grouped=True, graph=False)
    sim.set_tuning(8, 2, 0)
    sim.exec(Ap, [8000] * 24, Bp, [8000] * 24, Cp, [8000] * 24)
    sim.free()
    del dA8, dB8, dC8, Afull
    # the FP64 Gram of a 2e6 x 512 operand
    dG = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (512, 2000000)).T)
    dS = hip.DeviceArray.zeros((512, 512))
    hip.gram(512, 2000000, dG, dS)
    hip.sync()
    emit({"suite": "pmc-calib", "copy_bytes_read": nel * 8, "copy_bytes_written": nel * 8,
          "dgemm_n": n, "dgemm_cases": "NN,NT,TN,TT (direct-to-LDS stages), then NN register-staged", "fold": "3 x 2000x1334, 3 x 8000x8000",
          "pack_elements": sum(d[4] * d[5] for d in pack), "ozaki": f"{n}^3 s=7",
          "ozaki_crt": f"{n}^3 moduli=15", "grid": "8000^3 on 4x2x3: per-rank exec, then grouped exec",
          "gram": "512 x 512 from 2e6 x 512"})


def dgemm_group(args):
    """hip_dgemm at n^3 for each tile-row group size (CA3DMM_DGEMM_GROUP =
    4, 8, 16, 32, 64): one untimed and three timed launches per trans case
    (the counter form of the sweep: one timed launch, --steps 1)."""
    n = args.n
    rnd = np.random.default_rng(3)
    dA = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n, n)))
    dB = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n, n)))
    dC = hip.DeviceArray.zeros((n, n))
    hip.set_sync(False)
    reps = max(1, min(args.steps, 3))
    for G in (4, 8, 16, 32, 64):
        os.environ["CA3DMM_DGEMM_GROUP"] = str(G)
        out = {}
        for ta, tb in ((0, 0), (0, 1), (1, 0), (1, 1)):
            ms = timed(lambda: hip.dgemm(ta, tb, n, n, n, 1.0, dA, dB, 0.0, dC), reps=reps)
            out[f"{'NT'[ta]}{'NT'[tb]}"] = round(2.0 * n ** 3 / ms / 1e9, 2)
        emit({"suite": "dgemm-group", "n": n, "group": G, "tflops": out})
    del os.environ["CA3DMM_DGEMM_GROUP"]
    hip.set_sync(True)


def grid_batch(args):
    """The simulated 4x2x3 grid (example_AB's 1D column layouts, NN) on
    8000^3 and 10007^3 in its per-rank mode, whose Cannon rounds run as one
    dgemm_batch_kernel launch: one untimed and --steps (at most 3) timed
    execs each (the counter form: --steps 1)."""
    rnd = np.random.default_rng(5)
    hip.set_sync(False)
    for n in (8000, 10007):
        p = 24
        lay = col1d(n, n, p)
        dA = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n, n)))
        dB = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n, n)))
        dC = hip.DeviceArray.zeros((n, n))
        Ap = [dA.at(r[2] * n) for r in lay]
        Bp = [dB.at(r[2] * n) for r in lay]
        Cp = [dC.at(r[2] * n) for r in lay]
        sim = E.SimGrid(p, n, n, n, 0, 0, lay, lay, lay, proc_grid=[4, 2, 3])
        ms = timed(lambda: sim.exec(Ap, [n] * p, Bp, [n] * p, Cp, [n] * p), reps=max(1, min(args.steps, 3)))
        emit({"suite": "grid-batch", "n": n, "grid": "4x2x3", "ms": round(ms, 3),
              "tflops": round(2.0 * n ** 3 / ms / 1e9, 2)})
        sim.free()
        del dA, dB, dC
    hip.set_sync(True)


class _At:
    """A device address inside a DeviceArray, as an operand of hip.dgemm / hip.gram."""
    def __init__(self, ptr):
        self.ptr = ptr


def dgemm_loads(args):
    """The direct-to-LDS load forms (CA3DMM_DGEMM_LOADS unset = 16-byte
    loads from any 8-byte aligned operand, "aligned" = 16-byte loads only
    for 16-byte aligned operands with even leading dimensions, "reg" =
    register staging) on operands with odd leading dimensions or odd
    offsets: rank 0's two Cannon rounds of the 8000^3 4x2x3 grid
    (2000 x 4000 x (1334 + 1333), the second with ldb = 1333), n^3 in the
    four trans cases with lda = ldb = n + 1, n^3 NN with both operands one
    double past a 16-byte boundary, and the 512 x 512 Gram of a 1000001 x
    512 operand.  One untimed and three timed calls each (the rank-share
    rounds: 20 timed calls, three interleaved passes, the best kept)."""
    n = args.n
    rnd = np.random.default_rng(4)
    hip.set_sync(False)
    m, nn, k0, k1 = 2000, 4000, 1334, 1333
    A0 = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (m, k0)))
    A1 = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (m, k1)))
    B0 = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (k0, nn)))
    B1 = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (k1, nn)))
    Cr = hip.DeviceArray.zeros((m, nn))
    Ao = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n + 1, n)))
    Bo = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n + 1, n)))
    Cn = hip.DeviceArray.zeros((n, n))
    kg = 1000001
    G = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (kg, 512)))
    S = hip.DeviceArray.zeros((512, 512))
    def use(mode):
        if mode is None:
            os.environ.pop("CA3DMM_DGEMM_LOADS", None)
        else:
            os.environ["CA3DMM_DGEMM_LOADS"] = mode

    def rounds():
        hip.dgemm(0, 0, m, nn, k0, 1.0, A0, B0, 0.0, Cr)
        hip.dgemm(0, 0, m, nn, k1, 1.0, A1, B1, 1.0, Cr)
    # the rank-share rounds, short launches: the modes interleaved, three
    # passes of 20 timed calls each, best pass reported
    best = {}
    for _ in range(3):
        for mode in (None, "aligned", "reg"):
            use(mode)
            r0 = timed(lambda: hip.dgemm(0, 0, m, nn, k0, 1.0, A0, B0, 0.0, Cr), reps=20)
            r1 = timed(lambda: hip.dgemm(0, 0, m, nn, k1, 1.0, A1, B1, 1.0, Cr), reps=20)
            both = timed(rounds, reps=20)
            b = best.setdefault(mode, [r0, r1, both])
            best[mode] = [min(x, y) for x, y in zip(b, (r0, r1, both))]
    for mode in (None, "aligned", "reg"):
        r0, r1, both = best[mode]
        emit({"suite": "dgemm-loads", "loads": mode or "16-byte", "rank0_round0_ldb1334_tflops":
              round(2.0 * m * nn * k0 / r0 / 1e9, 2), "rank0_round1_ldb1333_tflops": round(2.0 * m * nn * k1 / r1 / 1e9, 2),
              "rank0_both_rounds_tflops": round(2.0 * m * nn * (k0 + k1) / both / 1e9, 2)})
    for mode in (None, "aligned", "reg"):
        use(mode)
        out = {}
        for ta, tb in ((0, 0), (0, 1), (1, 0), (1, 1)):
            ms = timed(lambda: hip.dgemm(ta, tb, n, n, n, 1.0, Ao, Bo, 0.0, Cn, lda=n + 1, ldb=n + 1), reps=3)
            out[f"ld_odd_{'NT'[ta]}{'NT'[tb]}"] = round(2.0 * n ** 3 / ms / 1e9, 2)
        ms = timed(lambda: hip.dgemm(0, 0, n, n, n, 1.0, _At(Ao.at(1)), _At(Bo.at(1)), 0.0, Cn, lda=n + 1,
                                     ldb=n + 1), reps=3)
        out["offset_1_ld_odd_NN"] = round(2.0 * n ** 3 / ms / 1e9, 2)
        ms = timed(lambda: hip.gram(512, kg, G, S), reps=3)
        out["gram_512_from_1000001"] = round(513.0 * 512 * kg / ms / 1e9, 2)
        emit({"suite": "dgemm-loads", "n": n, "loads": mode or "16-byte", "tflops": out})
    os.environ.pop("CA3DMM_DGEMM_LOADS", None)
    hip.set_sync(True)


def yardstick(args):
    """rocBLAS dgemm and INT8 gemm_ex at n^3, timed as a yardstick for the
    hand-written kernels (the product path never calls rocBLAS)."""
    import ctypes as C
    rb = C.CDLL("librocblas.so")
    h = C.c_void_p()
    assert rb.rocblas_create_handle(C.byref(h)) == 0
    rb.rocblas_dgemm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(C.c_double), C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                 C.POINTER(C.c_double), C.c_void_p, C.c_int]
    rb.rocblas_gemm_ex.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                   C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                   C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int32, C.c_uint32]
    hiprt = C.CDLL("libamdhip64.so")
    dsync = lambda: hiprt.hipDeviceSynchronize()
    NONE, TRANS, F64, I8, I32 = 111, 112, 152, 160, 162
    n = args.n
    rnd = np.random.default_rng(9)
    dA = hip.DeviceArray.from_numpy(rnd.uniform(-1, 1, (n, n)))
    dC = hip.DeviceArray.zeros((n, n))
    one, zero = C.c_double(1.0), C.c_double(0.0)

    def wall(f, reps):
        f()
        dsync()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        dsync()
        return (time.perf_counter() - t0) * 1e3 / reps
    for ta, tb, name in ((NONE, NONE, "NN"), (TRANS, NONE, "TN")):
        ms = wall(lambda: rb.rocblas_dgemm(h, ta, tb, n, n, n, C.byref(one), dA.ptr, n, dA.ptr, n,
                                           C.byref(zero), dC.ptr, n), 5)
        emit({"suite": "yardstick", "lib": "rocblas_dgemm", "case": name, "n": n, "ms": round(ms, 3),
              "tflops": round(2.0 * n ** 3 / ms / 1e9, 2)})
    del dA, dC
    a8 = hip.DeviceArray.filled((n, n), 1, dtype=np.int8)
    c32 = hip.DeviceArray.zeros((n, n), dtype=np.int32)
    ia, ib = C.c_int32(1), C.c_int32(0)
    ms = wall(lambda: rb.rocblas_gemm_ex(h, TRANS, NONE, n, n, n, C.byref(ia), a8.ptr, I8, n, a8.ptr, I8, n,
                                         C.byref(ib), c32.ptr, I32, n, c32.ptr, I32, n, I32, 0, 0, 0), 5)
    emit({"suite": "yardstick", "lib": "rocblas_gemm_ex i8->i32", "case": "TN", "n": n, "ms": round(ms, 3),
          "tops": round(2.0 * n ** 3 / ms / 1e9, 1)})
    rb.rocblas_destroy_handle(h)


def run(name, args):
    {"yardstick": yardstick, "rank-share": rank_share, "ozaki-sweep": ozaki_sweep, "ozaki-forms": ozaki_forms, "grouped": grouped, "cholqr": cholqr,
     "pmc-calib": pmc_calib, "dgemm-group": dgemm_group, "dgemm-loads": dgemm_loads, "grid-batch": grid_batch}[name](args)
