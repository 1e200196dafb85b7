"""Helpers for running the engines on device buffers in tests (not a test module)."""
import ctypes as C

import numpy as np

from ca3dmm_amd import device as hip
from ca3dmm_amd import engine as E
from oracle import oracle as O


def linear_fill(nrow, ncol, a, b):
    """examples/example_AB.c:107-128: X(i,j) = a*i + b*j on stored indices."""
    i = np.arange(nrow)[:, None]
    j = np.arange(ncol)[None, :]
    return np.asfortranarray(a * i + b * j)


def run_sim(p, m, n, k, ta, tb, A, B, proc_grid=None, is_btb=False, keep=False, grouped=False,
            graph=False, autotune=False, ozaki=0, repeat=1):
    """Run the simulated p-rank grid with the example_AB 1D column layout.
    A, B: full matrices as stored (numpy, any order).  Returns the full C
    gathered from the ranks' dst_C blocks.  keep: also returns the SimGrid,
    the plan and (Ap, ldA, Bp, ldB, Cp, ldC, dC, dA, dB); the raw addresses
    stay valid only while that tuple (which owns dA, dB, dC) is alive."""
    w = O.ca3dmm_world(m, n, k, ta, tb, p, proc_grid=proc_grid, is_btb=is_btb)
    mm = n if is_btb else m
    dA = hip.DeviceArray.from_numpy(A) if not is_btb else None
    dB = hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.filled((mm, n), 0xFF)
    sim = E.SimGrid(p, m, n, k, ta, tb, None if is_btb else w["src_A"], w["src_B"], w["dst_C"],
                    proc_grid=proc_grid, is_btb=is_btb, allocate=True)
    Ap = [dA.at(r_[2] * A.shape[0]) for r_ in w["src_A"]] if not is_btb else None
    ldA = [A.shape[0]] * p if not is_btb else None
    Bp = [dB.at(r_[2] * B.shape[0]) for r_ in w["src_B"]]
    ldB = [B.shape[0]] * p
    Cp = [dC.at(r_[2] * mm) for r_ in w["dst_C"]]
    ldC = [mm] * p
    if ozaki:
        sim.set_ozaki(ozaki)
    if grouped or graph:
        sim.set_mode(grouped=grouped, graph=graph)
        if autotune:
            sim.autotune(1)
    for _ in range(repeat):
        sim.exec(Ap, ldA, Bp, ldB, Cp, ldC)
    out = dC.numpy()
    if keep:
        return out, sim, w, (Ap, ldA, Bp, ldB, Cp, ldC, dC, dA, dB)
    sim.free()
    return out


def single_engine(m, n, k, ta, tb, dst=None, btb=False):
    eng = E.CA3P()
    An, Ac = (k, m) if ta else (m, k)
    Bn, Bc = (n, k) if tb else (k, n)
    d = dst if dst is not None else (0, m if not btb else n, 0, n)
    if btb:
        E.lib.ca3dmm_engine_init_BTB(n, k, 0, k, 0, n, d[0], d[1], d[2], d[3], None, E.comm_self(),
                                     E.DEV_TYPE_HIP, C.byref(eng), None)
    else:
        E.lib.ca3dmm_engine_init(m, n, k, ta, tb, 0, An, 0, Ac, 0, Bn, 0, Bc, d[0], d[1], d[2], d[3],
                                 None, E.comm_self(), E.DEV_TYPE_HIP, C.byref(eng), None)
    assert eng, "engine init failed"
    return eng


def ozaki_grid_reference(m, n, k, p, proc_grid, opA, opB, nmod, is_btb=False):
    """The Ozaki form of the simulated grid, restated: every C block of the
    2D grid is the sum over the k layers, in layer order, of the CRT-form
    product (oracle/ozaki.py, ozaki2_dgemm) of the layer's operand blocks
    (for B^T B: opA = B^T, opB = B on the B^T B grid)."""
    from oracle import ozaki as Z
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=proc_grid, is_btb=is_btb)
    out = np.full((m, n), np.nan)
    for pl in w["plans"]:
        if not pl.is_active or pl.rank_k != 0 or not pl.C_2dmm_nrow or not pl.C_2dmm_ncol:
            continue
        rs = slice(pl.C_2dmm_srow, pl.C_2dmm_srow + pl.C_2dmm_nrow)
        cs = slice(pl.C_2dmm_scol, pl.C_2dmm_scol + pl.C_2dmm_ncol)
        acc = None
        for layer in range(pl.task_k_num):
            k0, kn = O.block_spos_size(k, pl.task_k_num, layer)
            v = Z.ozaki2_dgemm(opA[rs, k0:k0 + kn], opB[k0:k0 + kn, cs], nmod)
            acc = v if acc is None else acc + v
        out[rs, cs] = acc
    return out
