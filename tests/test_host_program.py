"""The host program (lbm-dem_amd/bin/lbdem_run: the reference's main loop, configuration reader and
result files over the C ABI) against the result files the reference program itself writes for
the same case (tests/golden/host_<case>/, written by make_golden.py from a single-threaded run)."""
import os
import subprocess

import numpy as np
import pytest

import cases
from conftest import GPU_LIB, ORACLE_LIB

RUN = os.path.join(cases.REPO, "lbm-dem_amd", "bin", "lbdem_run")
GOLD = os.path.join(cases.REPO, "tests", "golden")


@pytest.fixture(scope="session")
def host_program():
    if not os.path.exists(RUN):
        subprocess.run(["make", "-C", os.path.join(cases.REPO, "lbm-dem_amd"), "host"], check=True,
                       capture_output=True)
    return RUN


def run_case(prog, root, name, library, extra=(), overrides=None):
    over = dict(cases.HOST_RUNS[name])
    over.update(overrides or {})
    d = cases.make_case(str(root), name, over)
    r = subprocess.run([prog, "-c", "config.cfg", "-d", "out", "--library", library, *extra], cwd=d,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return os.path.join(d, "out"), r.stdout


def table(path):
    with open(path) as fh:
        lines = fh.read().splitlines()[1:]
    return np.array([[float(t) for t in l.split()] for l in lines])


@pytest.mark.parametrize("name", sorted(cases.HOST_RUNS))
def test_host_program_writes_the_reference_result_files(host_program, oracle_lib, tmp_path, name):
    """With the FP64 oracle behind the ABI the files are byte-identical to the reference's."""
    out, _ = run_case(host_program, tmp_path, name, ORACLE_LIB)
    for f in cases.HOST_FILES:
        with open(os.path.join(out, f)) as a, open(os.path.join(GOLD, f"host_{name}", f)) as b:
            assert a.read() == b.read(), f"{name}: {f}"


def test_command_line_overrides_the_configuration(host_program, oracle_lib, tmp_path):
    """`-name value` replaces the file's value (macros.h:14-22): a shorter time window gives fewer
    rows, and the first rows equal the full run's."""
    out, log = run_case(host_program, tmp_path, "suspension", ORACLE_LIB, extra=("-maxTime", "11"))
    pos = table(os.path.join(out, "objPosition.dat"))
    ref = table(os.path.join(GOLD, "host_suspension", "objPosition.dat"))
    assert len(pos) == 4  # realTime 0, 4, 8 and the end of step 12 is past maxTime
    assert np.array_equal(pos, ref[:4])
    mass = table(os.path.join(out, "mass.dat"))
    assert mass.shape == (4, 3) and np.all(mass[:, 2] > 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(cases.HOST_RUNS))
def test_host_program_on_the_gpu(host_program, gpu_lib, tmp_path, name):
    """The MI355X library behind the same host program (FP64): the reference's result files to
    the printed precision (6 significant digits; values below 1e-9 of the column's scale are
    round-off and compared absolutely)."""
    out, _ = run_case(host_program, tmp_path, name, GPU_LIB, extra=("--precision", "fp64"))
    for f in cases.HOST_FILES:
        a = table(os.path.join(out, f))
        b = table(os.path.join(GOLD, f"host_{name}", f))
        assert a.shape == b.shape, f
        if b.size == 0:  # a header only (bodyPosition.dat of a run without bodies)
            continue
        scale = np.maximum(np.abs(b).max(0), 1e-30)
        err = np.abs(a - b) / scale
        assert err.max() <= 2e-5, f"{name}: {f} differs by {err.max()} of the column scale"


# ---------------------------------------------------------------------------------------------
# fluid restart files (IO.cpp:973-997 writes them, LB.cpp:352-367 reads them back)
# ---------------------------------------------------------------------------------------------
def recycle_text(step):
    import gzip
    with gzip.open(os.path.join(GOLD, "recycle_suspension", cases.recycle_name(step) + ".gz"), "rt") as fh:
        return fh.read()


def gold_text(folder, fname):
    import gzip
    with gzip.open(os.path.join(GOLD, folder, fname + ".gz"), "rt") as fh:
        return fh.read()


def run_recycle(prog, root, library, extra=(), run=None):
    name, over = run or cases.RECYCLE_RUN
    d = cases.make_case(str(root), name, over)
    r = subprocess.run([prog, "-c", "config.cfg", "-d", "out", "--library", library, *extra], cwd=d,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return os.path.join(d, "out", "fluidData")


def recycle_table(text):
    """(size, types, records) of a fluidRecycle file's text, records as rows of number strings."""
    lines = text.splitlines()
    size = [int(t) for t in lines[0].split()]
    types = np.array([int(t) for t in lines[1].split()], np.int64)
    return size, types, [l.split() for l in lines[2:]]


def test_host_program_writes_the_reference_recycle_files(host_program, oracle_lib, tmp_path):
    """fluidRecycleExpTime: with the FP64 oracle behind the ABI the fluidRecycle files are
    byte-identical to the reference's, at steps 0, 15 and 30 (every node's type, then per active
    node n - 1, u, mass - 1, viscosity and the 19 populations)."""
    out = run_recycle(host_program, tmp_path, ORACLE_LIB)
    assert sorted(os.listdir(out)) == [cases.recycle_name(s) for s in cases.RECYCLE_STEPS]
    for s in cases.RECYCLE_STEPS:
        with open(os.path.join(out, cases.recycle_name(s))) as fh:
            assert fh.read() == recycle_text(s), f"step {s}"


def restarted_file(prog, root, library, extra=()):
    """The recycle file the host program writes right after restarting from the reference's
    step-15 file (the restart case, one step, a recycle file at step 0 only)."""
    d = cases.make_case(str(root), "restart", {"maxTime": 0.0, "fluidRecycleExpTime": 100.0})
    r = subprocess.run([prog, "-c", "config.cfg", "-d", "out", "--library", library, *extra], cwd=d,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    with open(os.path.join(d, "out", "fluidData", cases.recycle_name(0))) as fh:
        return recycle_table(fh.read())


def test_host_program_restarts_from_the_reference_file(host_program, oracle_lib, tmp_path):
    """lbRestart: the host program parses the reference's file and restarts from it; what it then
    writes holds the reference's restarted types (tests/golden/restart_s0, from the reference's own
    restart) and, as the reference prints them, its densities, masses and populations."""
    size, types, rows = restarted_file(host_program, tmp_path, ORACLE_LIB)
    g = np.load(os.path.join(GOLD, "restart_s0.npz"), allow_pickle=False)
    assert size == list(g["size"]) and np.array_equal(types, g["type"])
    act = g["active"]
    assert len(rows) == len(act)
    fmt = np.vectorize(lambda v: "%g" % v)
    assert [r[0] for r in rows] == list(fmt(g["n_active"] - 1.0))
    assert [r[4] for r in rows] == list(fmt(g["mass_active"] - 1.0))
    for q in range(19):
        assert [r[6 + q] for r in rows] == list(fmt(g["f_active"][q])), f"population {q}"


@pytest.mark.gpu
def test_host_program_recycle_files_on_the_gpu(host_program, gpu_lib, tmp_path):
    """The MI355X library behind the host program (FP64) writes the reference's fluidRecycle
    files: the same types and records, the numbers to the printed precision (relative to the
    column's scale), and restarts from the reference's file to its restarted state."""
    out = run_recycle(host_program, tmp_path / "w", GPU_LIB, extra=("--precision", "fp64"))
    for s in cases.RECYCLE_STEPS:
        with open(os.path.join(out, cases.recycle_name(s))) as fh:
            size, types, rows = recycle_table(fh.read())
        gs, gt, grows = recycle_table(recycle_text(s))
        assert size == gs and np.array_equal(types, gt) and len(rows) == len(grows), f"step {s}"
        a = np.array(rows, np.float64)
        b = np.array(grows, np.float64)
        scale = np.maximum(np.abs(b).max(0), 1e-30)
        assert (np.abs(a - b) / scale).max() <= 2e-5, f"step {s}"
    size, types, rows = restarted_file(host_program, tmp_path / "r", GPU_LIB, extra=("--precision", "fp64"))
    g = np.load(os.path.join(GOLD, "restart_s0.npz"), allow_pickle=False)
    assert np.array_equal(types, g["type"])
    a = np.array(rows, np.float64)
    assert np.abs(a[:, 6:] - g["f_active"].T).max() <= 1e-6 * np.abs(g["f_active"]).max()
    assert np.abs(a[:, 4] - (g["mass_active"] - 1.0)).max() <= 1e-6


# ---------------------------------------------------------------------------------------------
# Eulerian (.vti) and Lagrangian (.vtu) fluid files (IO.cpp:186-202, 1038-1235)
# ---------------------------------------------------------------------------------------------
FLUID_DIR = "fluid_" + cases.FLUID_RUN[0]


def fluid_gold_names():
    return sorted(f[:-3] for f in os.listdir(os.path.join(GOLD, FLUID_DIR)))


def test_host_program_writes_the_reference_fluid_files(host_program, oracle_lib, tmp_path):
    """fluidExpTime / fluidLagrangianExpTime in physical units (the units case): with the FP64
    oracle behind the ABI the .vti (every node's type and physical mass) and .vtu files (the active
    nodes' velocity, pressure, density, viscosity, mass, type and position) are byte-identical to
    the reference's."""
    out = run_recycle(host_program, tmp_path, ORACLE_LIB, run=cases.FLUID_RUN)
    assert sorted(os.listdir(out)) == fluid_gold_names()
    for f in fluid_gold_names():
        with open(os.path.join(out, f)) as fh:
            assert fh.read() == gold_text(FLUID_DIR, f), f


@pytest.mark.gpu
def test_host_program_fluid_files_on_the_gpu(host_program, gpu_lib, tmp_path):
    """The MI355X library behind the host program (FP64) writes the reference's .vti and .vtu
    files: the same tags, node sets and types, the numbers to the printed precision."""
    import re
    out = run_recycle(host_program, tmp_path, GPU_LIB, extra=("--precision", "fp64"), run=cases.FLUID_RUN)
    assert sorted(os.listdir(out)) == fluid_gold_names()
    for f in fluid_gold_names():
        with open(os.path.join(out, f)) as fh:
            a_text = fh.read()
        b_text = gold_text(FLUID_DIR, f)
        assert re.findall(r"<[^>]*>", a_text) == re.findall(r"<[^>]*>", b_text), f
        # section by section (the text between tags: one data array each), relative to its scale
        for k, (sa, sb) in enumerate(zip(re.split(r"<[^>]*>", a_text), re.split(r"<[^>]*>", b_text))):
            a = np.array([float(t) for t in sa.split()], np.float64)
            b = np.array([float(t) for t in sb.split()], np.float64)
            assert a.shape == b.shape, f"{f} section {k}"
            if b.size:
                assert np.abs(a - b).max() <= 2e-5 * max(np.abs(b).max(), 1e-30), f"{f} section {k}"


# ---------------------------------------------------------------------------------------------
# DEM VTK files (IO.cpp:212-256, 651-912): particle velocities, box edges, impactors, bodies and
# springs, every one on partExpTime
# ---------------------------------------------------------------------------------------------
DEM_DIR = "dem_" + cases.DEM_RUN[0]


def dem_gold_names():
    return sorted(f[:-3] for f in os.listdir(os.path.join(GOLD, DEM_DIR)))


def run_dem(prog, root, library, extra=()):
    name, over = cases.DEM_RUN
    d = cases.make_case(str(root), name, over)
    r = subprocess.run([prog, "-c", "config.cfg", "-d", "out", "--library", library, *extra], cwd=d,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    out = os.path.join(d, "out", "particleData")
    return out, sorted(f for f in os.listdir(out) if f.startswith(cases.DEM_FILES))


def test_host_program_writes_the_reference_dem_files(host_program, oracle_lib, tmp_path):
    """With the FP64 oracle behind the ABI the velocity, boundary, impactor, body and spring VTK
    files are byte-identical to the reference's (fixed 8-digit coordinates, the box of the fluid
    height, the spring file's two points and compressions as the reference writes them)."""
    out, names = run_dem(host_program, tmp_path, ORACLE_LIB)
    assert names == dem_gold_names()
    for f in names:
        with open(os.path.join(out, f)) as fh:
            assert fh.read() == gold_text(DEM_DIR, f), f


@pytest.mark.gpu
def test_host_program_dem_files_on_the_gpu(host_program, gpu_lib, tmp_path):
    """The MI355X library behind the host program (FP64) writes the same DEM VTK files: the same
    words, the numbers to the printed precision."""
    out, names = run_dem(host_program, tmp_path, GPU_LIB, extra=("--precision", "fp64"))
    assert names == dem_gold_names()
    for f in names:
        with open(os.path.join(out, f)) as fh:
            a_tok = fh.read().split()
        b_tok = gold_text(DEM_DIR, f).split()
        assert len(a_tok) == len(b_tok), f

        def num(t):
            try:
                return float(t)
            except ValueError:
                return None
        an, bn = [num(t) for t in a_tok], [num(t) for t in b_tok]
        assert [t for t, v in zip(a_tok, an) if v is None] == [t for t, v in zip(b_tok, bn) if v is None], f
        a = np.array([v for v in an if v is not None])
        b = np.array([v for v in bn if v is not None])
        if b.size:
            assert np.abs(a - b).max() <= 1e-6 * max(np.abs(b).max(), 1.0), f
