"""GPU: every measured configuration at its own size, against the oracle.

  C2  8192^3 FP64 DGEMM, all four transA/transB cases
      (cuda_cublas_dgemm replacement, src/cuda_proxy.cu:211-230), and the
      8000^3 problem on the 4x2x3 grid of 24 ranks (README.md:58-98): rank 0's
      cuboid 2000 x 4000 x (1334 + 1333), the kp = 3 partial-C sum and the
      pack / unpack of the 1D column layout (examples/example_AB.c:85-99);
  C3  the Ozaki forms at 8192^3 (phi 0.5, 2, 4) and 65536 x 65536 x 1024
      (phi 0.5) at the counts hip_ozaki_choose picks;
  C4  10007^3 on the simulated 4x2x3 grid in one persistent grouped launch,
      replayed from a hipGraph;
  C5  CholeskyQR2 and shifted CholeskyQR3 of 2e6 x 512 at kappa 1e2, 1e5, 1e8.

Floating-point bars: |C - C_ref| <= k u (|A||B|)_ij componentwise (u = 2^-53)
against double-double dot products on 4096 sampled entries (the validation of
examples/example_AB.c:287-305 at the tolerance its BLAS allows); the Ozaki forms
no worse than the FP64 DGEMM in both max |dC| / (|A||B|) and max relative
error; CholeskyQR ||Q^T Q - I||_F < 1e-13 and ||A - QR||_F / ||A||_F < 1e-14.
Integer-valued inputs must come out exact (every summation order is exact)."""
import sys

import numpy as np
import pytest

from ca3dmm_amd import cholqr
from oracle import oracle as O
from oracle import ozaki as Z
from tests.simutil import run_sim

pytestmark = pytest.mark.gpu
U = 2.0 ** -53
SAMPLES = 4096


def say(msg):
    print(f"[configs] {msg}", flush=True)
    sys.stdout.flush()


def fmat(rnd, nrow, ncol, lo=-1.0, hi=1.0):
    """uniform[lo, hi] in column-major order without a copy."""
    return rnd.uniform(lo, hi, (ncol, nrow)).T


def samples(rnd, m, n, s=SAMPLES):
    return rnd.integers(0, m, s).astype(np.int32), rnd.integers(0, n, s).astype(np.int32)


def errors(got, ref, ab):
    d = np.abs(got - ref)
    return float(np.max(d / ab)), float(np.max(d / np.abs(ref)))


# ---------------------------------------------------------------------------- C2
def test_c2_dgemm_8192_all_trans(hip):
    m = n = k = 8192
    rnd = np.random.default_rng(2022)
    A, B = fmat(rnd, m, k), fmat(rnd, k, n)          # square: also serve as the transposed operands
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.filled((m, n), 0xFF)
    ii, jj = samples(rnd, m, n)
    for ta, tb in ((0, 0), (0, 1), (1, 0), (1, 1)):
        say(f"C2 dgemm 8192^3 transA={ta} transB={tb}")
        hip.dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, dC)
        got = dC.sample(ii, jj)
        ref, ab = O.dd_sample(ta, tb, m, n, k, A, m, B, k, ii, jj)
        err = np.abs(got - ref)
        assert np.all(err <= k * U * ab), (ta, tb, float(np.max(err / ab)))
        say(f"  max |dC|/(|A||B|) = {float(np.max(err / ab)):.3e}")


def test_c2_rank_share_8000_grid(hip):
    """8000^3 NN on the 4x2x3 grid (p = 24), 1D column layout of A, B and C:
    the redistributions, both Cannon rounds of every rank, the kp = 3 fold
    and the C redistribution, per-launch and in the grouped launch."""
    m = n = k = 8000
    rnd = np.random.default_rng(24)
    A, B = fmat(rnd, m, k), fmat(rnd, k, n)
    ii, jj = samples(rnd, m, n)
    ref, ab = O.dd_sample(0, 0, m, n, k, A, m, B, k, ii, jj)
    for grouped in (False, True):
        say(f"C2 rank share 8000^3 on 4x2x3, grouped={grouped}")
        got, sim, w, args = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=grouped, keep=True)
        e0 = sim.engine(0)
        assert (e0.mp, e0.np, e0.kp) == (4, 2, 3)
        assert (e0.C_2dmm_nrow, e0.C_2dmm_ncol) == (2000, 4000)
        assert (e0.A_2dmm_nrow, e0.A_2dmm_ncol) == (2000, 1334)      # Cannon block: round 0's k-block
        ce = e0.cannon_engine.contents
        assert ce.np_dim == 2 and [ce.k_displs[i] for i in range(3)] == [0, 1334, 2667]
        assert (e0.C_out_nrow, e0.C_out_ncol) == (2000, 1334)       # owner slice of the kp = 3 fold
        sim.free()
        del args
        err = np.abs(got[ii, jj] - ref)
        assert np.all(err <= k * U * ab), float(np.max(err / ab))
        assert np.all(np.isfinite(got))                              # every entry written (dst was NaN)
        say(f"  max |dC|/(|A||B|) = {float(np.max(err / ab)):.3e}")


# ---------------------------------------------------------------------------- C3
def _ozaki_case(hip, m, n, k, phi, seed, forms):
    rnd = np.random.default_rng(seed)
    A = np.asfortranarray(Z.phi_matrix(rnd, (m, k), phi))
    B = np.asfortranarray(Z.phi_matrix(rnd, (k, n), phi))
    ii, jj = samples(rnd, m, n)
    ref, ab = O.dd_sample(0, 0, m, n, k, A, m, B, k, ii, jj)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    del A, B
    dC = hip.DeviceArray((m, n))
    hip.dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC)
    nat = errors(dC.sample(ii, jj), ref, ab)
    s, N = hip.ozaki_choose(0, 0, m, n, k, dA, dB)
    say(f"C3 {m}x{n}x{k} phi={phi}: native {nat[0]:.3e} / {nat[1]:.3e}, chooser s={s} N={N}")
    out = {}
    if "crt" in forms:
        hip.ozaki2_dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC, N)
        out["crt"] = errors(dC.sample(ii, jj), ref, ab)
    if "slice" in forms:
        hip.ozaki_dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC, s)
        out["slice"] = errors(dC.sample(ii, jj), ref, ab)
    for f, e in out.items():
        say(f"  {f}: {e[0]:.3e} / {e[1]:.3e}")
        assert e[0] <= nat[0] and e[1] <= nat[1], (f, e, nat)


@pytest.mark.parametrize("phi", [0.5, 2.0, 4.0])
def test_c3_ozaki_8192_at_chosen_counts(hip, phi):
    _ozaki_case(hip, 8192, 8192, 8192, phi, 3000 + int(10 * phi), ("crt", "slice"))


def test_c3_ozaki_65536x65536x1024(hip):
    _ozaki_case(hip, 65536, 65536, 1024, 0.5, 3100, ("crt",))


# ---------------------------------------------------------------------------- C4
def test_c4_grouped_10007_graph(hip):
    """10007^3 on 4x2x3: 24 uneven cuboids, q = 2 Cannon rounds and the kp = 3
    fold in one persistent launch, autotuned, replayed from a hipGraph."""
    from ca3dmm_amd import device as dev
    m = n = k = 10007
    rnd = np.random.default_rng(10007)
    # integer-valued inputs: exact in any summation order
    Ai = np.asfortranarray(rnd.integers(-4, 5, (m, k)).astype(np.float64))
    Bi = np.asfortranarray(rnd.integers(-4, 5, (k, n)).astype(np.float64))
    say("C4 10007^3 grouped + hipGraph, integer inputs")
    got = run_sim(24, m, n, k, 0, 0, Ai, Bi, proc_grid=[4, 2, 3], grouped=True, graph=True, autotune=True)
    assert np.array_equal(got, Ai @ Bi)
    del Ai, Bi, got
    A, B = fmat(rnd, m, k), fmat(rnd, k, n)
    ii, jj = samples(rnd, m, n)
    ref, ab = O.dd_sample(0, 0, m, n, k, A, m, B, k, ii, jj)
    say("C4 10007^3 grouped eager, uniform inputs")
    eager = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True, graph=False)
    err = np.abs(eager[ii, jj] - ref)
    assert np.all(err <= k * U * ab), float(np.max(err / ab))
    say("C4 10007^3 grouped + hipGraph replays")
    got, sim, w, args = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True, graph=True,
                                keep=True)
    Ap, ldA, Bp, ldB, Cp, ldC, dC, dA, dB = args
    assert np.array_equal(got, eager)
    for rep in range(3):
        dev.sync()
        sim.exec(Ap, ldA, Bp, ldB, Cp, ldC)
        assert np.array_equal(dC.numpy(), eager), rep
    sim.free()


# ---------------------------------------------------------------------------- C5
@pytest.fixture(scope="module")
def tall_u(hip):
    m, n = 2000000, 512
    rnd = np.random.default_rng(2022)
    say("C5 building U: CholeskyQR2 of a 2e6 x 512 Gaussian")
    dG = hip.DeviceArray.from_numpy(rnd.standard_normal((n, m)).T)
    dU, _ = cholqr.cholesky_qr2(dG, m, n)
    del dG
    yield dU


@pytest.mark.parametrize("cond", [1e2, 1e5, 1e8])
def test_c5_cholesky_qr_2e6x512(hip, tall_u, cond):
    m, n = 2000000, 512
    dA, _ = cholqr.test_matrix(m, n, cond, seed=2022, dU=tall_u)
    cases = [(hip.CHOLQR2, None), (hip.CHOLQR2, 8)]
    if cond >= 1e8:
        cases += [(hip.SCHOLQR3, None), (hip.CHOLQR_AUTO, None)]
    for variant, gs in cases:
        dQ, dR, passes = cholqr.cholesky_qr(dA, m, n, variant, gram_slices=gs)
        orth, res = cholqr.qr_errors(dA, dQ, dR, m, n)
        say(f"C5 kappa={cond:g} variant={variant} gram_slices={gs}: passes {passes}, "
            f"||Q^TQ-I||_F {orth:.3e}, ||A-QR||/||A|| {res:.3e}")
        assert passes == (3 if variant == hip.SCHOLQR3 else 2)
        assert orth < 1e-13, orth
        assert res < 1e-14, res
        del dQ, dR
