"""lbdem_get_fields_noalloc (downloads through the library's file-scope device scratch, no
allocation) against lbdem_get_fields on the same state: identical populations, mass and density
fields on two calls in a row, the second on a smaller box, for each storage precision; a box of
more than 2^21 nodes is rejected with LBDEM_E_CAPACITY before any device work."""
import numpy as np
import pytest

import cases
import lbdem

pytestmark = pytest.mark.gpu


def _same_fields(sim, where):
    f0, m0, n0 = sim.fields(True)
    f1, m1, n1 = sim.fields(True, noalloc=True)
    assert np.array_equal(f0, f1), f"{where}: populations"
    assert np.array_equal(m0, m1), f"{where}: mass"
    assert np.array_equal(n0, n1), f"{where}: density"
    # the scratch keeps the last call's contents: a call that wants only the mass field still
    # returns the allocating entry point's values
    m2 = np.zeros_like(m0)
    sim._check(sim.lib.lib.lbdem_get_fields_noalloc(sim.ctx, None, m2.ctypes.data, None), "noalloc mass only")
    assert np.array_equal(m0, m2), f"{where}: mass alone"
    return float(np.abs(f0).sum())


@pytest.mark.parametrize("prec", [lbdem.PREC_FP64, lbdem.PREC_FP32, lbdem.PREC_FP16])
def test_noalloc_fields_equal_allocating_fields(gpu_lib, case_root, prec):
    big = cases.open_sim(gpu_lib, cases.make_case(case_root, "suspension"), precision=prec)
    small = cases.open_sim(gpu_lib, cases.make_case(case_root, "units"), precision=prec)
    assert small.info()["totNodes"] < big.info()["totNodes"]
    print(f"prec={prec} boxes {big.info()['size']} then {small.info()['size']}", flush=True)
    big.cycle(12)
    small.cycle(7)
    assert _same_fields(big, "suspension") > 0.0
    assert _same_fields(small, "units (smaller box, second call)") > 0.0
    big.close()
    small.close()


def test_noalloc_rejects_boxes_beyond_its_scratch(gpu_lib, case_root):
    vals, *_ = cases._case_values("lb_closed")
    vals.update(lbSizeX=130.0, lbSizeY=128.0, lbSizeZ=128.0)  # 2,129,920 nodes > 2^21
    prm = lbdem.params_from_values(vals, lbdem.PREC_FP32)
    sim = lbdem.Simulation(gpu_lib, prm, lbdem.ParticleSet(), lbdem.ObjectSet())
    sim.init()
    print("box", sim.info()["size"], sim.info()["totNodes"], flush=True)
    m = np.zeros(sim.info()["totNodes"])
    rc = gpu_lib.lib.lbdem_get_fields_noalloc(sim.ctx, None, m.ctypes.data, None)
    assert rc == 4  # LBDEM_E_CAPACITY
    assert b"2097152" in gpu_lib.lib.lbdem_last_error(sim.ctx)
    assert not m.any()
    # the allocating entry point still serves the box
    _, m0, _ = sim.fields(False)
    assert m0.sum() > 0.0
    sim.close()
