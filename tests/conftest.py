"""Shared fixtures: library loading (building them if absent) and the gpu marker."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "lbm-dem_amd", "python"))

ORACLE_LIB = os.path.join(REPO, "oracle", "_build", "liblbdem_refcpu.so")
GPU_LIB = os.path.join(REPO, "lbm-dem_amd", "lib", "liblbdem.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    import lbdem
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "all"], check=True, capture_output=True)
    return lbdem.Library.load(ORACLE_LIB)


@pytest.fixture(scope="session")
def gpu_lib():
    import lbdem
    if not os.path.exists(GPU_LIB):
        raise RuntimeError(f"{GPU_LIB} is not built: run __graft_entry__.build() first")
    return lbdem.Library.load(GPU_LIB)


@pytest.fixture(scope="session")
def case_root(tmp_path_factory):
    return str(tmp_path_factory.mktemp("cases"))
