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
        scale = np.maximum(np.abs(b).max(0), 1e-30)
        err = np.abs(a - b) / scale
        assert err.max() <= 2e-5, f"{name}: {f} differs by {err.max()} of the column scale"
