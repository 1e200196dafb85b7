"""CPU: the reference's MPI-typed constructors (include/ca3dmm_mpi.h over
libca3dmm_amd_mpi.so).  A caller written against src/ca3dmm.h:114-124,
:149-157, src/cannon.h:69-72 and src/mat_redist.h:69-74 -- MPI_Comm and
MPI_Datatype by value -- compiles with -DCA3DMM_MPI, links, and plans the same
engines as the ca3dmm_comm_t entries on one rank; a communicator of two ranks
is refused on every rank with a NULL engine (the reference's failure
convention, src/ca3dmm.c:36-40).  The example driver is built both ways."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ca3dmm_amd", "lib")
MPI_HOME = "/opt/conda"
HAVE_MPI = os.path.exists(os.path.join(MPI_HOME, "include", "mpi.h")) and os.path.exists(
    os.path.join(LIB, "libca3dmm_amd_mpi.so"))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    if not HAVE_MPI:
        pytest.skip("no MPI installation / libca3dmm_amd_mpi.so on this machine")
    out = str(tmp_path_factory.mktemp("mpi") / "mpi_boundary")
    libdir = os.path.dirname(os.path.realpath(subprocess.check_output(
        ["gcc", "-print-file-name=libstdc++.so"], text=True).strip()))
    cmd = ["gcc", "-O1", "-std=c11", "-Wall", "-Werror", "-DCA3DMM_MPI", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(MPI_HOME, "include"), os.path.join(ROOT, "tests", "c", "mpi_boundary.c"),
           "-o", out, "-L" + LIB, "-Wl,-rpath," + LIB, "-lca3dmm_amd_mpi", "-lca3dmm_amd",
           os.path.join(MPI_HOME, "lib", "libmpi.so"), "-Wl,-rpath-link," + libdir,
           "-Wl,--enable-new-dtags,-rpath," + os.path.join(MPI_HOME, "lib")]
    subprocess.run(cmd, check=True)
    return out


def _lines(text):
    return [l.split() for l in text.splitlines() if l.strip()]


def test_single_rank_plans_like_the_comm_entries(probe):
    r = subprocess.run([probe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = {l[0]: l for l in _lines(r.stdout)}
    for name in ("ca3dmm_engine_init", "ca3dmm_engine_init_BTB", "cannon_engine_init", "mat_redist_engine_init"):
        assert got[name][1:4] == ["0", "1", "ok"], (name, r.stdout, r.stderr)
    assert got["grid"][1:] == ["1", "1", "1", "rank", "0"]
    assert got["dtype"][1] == "0" and got["dtype"][3] == "8"
    # the same plan through the ca3dmm_comm_t entry point
    from ca3dmm_amd import engine as E
    eng, bytes_ = E.CA3P(), C.c_size_t()
    E.lib.ca3dmm_engine_init(300, 200, 500, 1, 0, 0, 500, 0, 300, 0, 500, 0, 200, 0, 300, 0, 200, None,
                             E.comm_self(), E.DEV_TYPE_HIP, C.byref(eng), C.byref(bytes_))
    assert eng
    assert int(got["ca3dmm_engine_init"][4]) == bytes_.value
    E.lib.ca3dmm_engine_free(C.byref(eng))
    rd, rbytes = C.POINTER(E.MatRedist)(), C.c_size_t()
    E.lib.mat_redist_engine_init(0, 0, 40, 30, 10, 5, 20, 20, E.comm_self(), 0, 8, E.DEV_TYPE_HIP, C.byref(rd),
                                 C.byref(rbytes))
    assert rd and int(got["mat_redist_engine_init"][4]) == rbytes.value
    E.lib.mat_redist_engine_free(C.byref(rd))


def test_two_ranks_are_refused(probe):
    mpiexec = os.path.join(MPI_HOME, "bin", "mpiexec")
    if not os.path.exists(mpiexec):
        pytest.skip("no mpiexec")
    r = subprocess.run([mpiexec, "-n", "2", probe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l for l in _lines(r.stdout) if l[0].endswith("_init") or l[0].endswith("_BTB")]
    assert len(rows) == 8
    assert all(l[2] == "2" and l[3] == "null" for l in rows), r.stdout
    assert "builds no inter-process exchange" in r.stderr


def test_example_drivers_are_built():
    exe = os.path.join(ROOT, "ca3dmm_amd", "examples", "bin", "example_AB")
    assert os.path.exists(exe), "build with `make -C ca3dmm_amd/examples` (done by __graft_entry__.build)"
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "transA(0 or 1)" in r.stdout
    if HAVE_MPI and shutil.which("gcc"):
        mexe = exe + "_mpi"
        assert os.path.exists(mexe)
        r = subprocess.run([mexe, "--help"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and "transA(0 or 1)" in r.stdout
