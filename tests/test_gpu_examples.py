"""GPU: the compiled example driver (ca3dmm_amd/examples/example_AB.c) -- the
reference's examples/example_AB.c command line, linear fill and validation
(an entry with relative error above 1e-12 counts as an error, :287-305) --
run at 2000^3 for all four transA/transB cases through ca3dmm_engine_exec on
the HIP device type, plain and built against MPI with the reference's
MPI-typed ca3dmm_engine_init (one rank, MPI_COMM_WORLD)."""
import os
import re
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ca3dmm_amd", "examples", "bin")


def _run(exe, args):
    print(f"[examples] {os.path.basename(exe)} {' '.join(args)}", flush=True)
    sys.stdout.flush()
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_example_ab_2000_zero_errors(hip, ta, tb):
    out = _run(os.path.join(BIN, "example_AB"), ["2000", "2000", "2000", str(ta), str(tb), "1", "5", "1"])
    assert "CA3DMM output : 0 error(s)" in out, out
    assert re.search(r"Process grid mp \* np \* kp\s+: 1 \* 1 \* 1", out), out
    assert "2D Cannon execution" in out                       # ca3dmm_engine_print_stat ran


def test_example_ab_uneven_shape_forced_redistribution(hip):
    """m, n, k not multiples of the tile; CA3DMM_FORCE_REDIST sends A, B and C
    through the pack / unpack path of the single-rank engine."""
    exe = os.path.join(BIN, "example_AB")
    env = dict(os.environ, CA3DMM_FORCE_REDIST="1")
    print("[examples] example_AB 1001 777 1999 1 0 with CA3DMM_FORCE_REDIST=1", flush=True)
    r = subprocess.run([exe, "1001", "777", "1999", "1", "0", "1", "2", "1"], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "CA3DMM output : 0 error(s)" in r.stdout, r.stdout


def test_example_ab_mpi_build_zero_errors(hip):
    exe = os.path.join(BIN, "example_AB_mpi")
    if not os.path.exists(exe):
        pytest.skip("example_AB_mpi not built (no MPI installation)")
    out = _run(exe, ["2000", "2000", "2000", "0", "0", "1", "3", "1"])
    assert "CA3DMM output : 0 error(s)" in out, out
