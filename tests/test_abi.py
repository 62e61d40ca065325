"""The C ABI: both libraries load and export every entry point include/lbdem.h declares, and the
host-side input readers follow the reference formats."""
import ctypes
import os
import re

import numpy as np
import pytest

import lbdem
from conftest import GPU_LIB, ORACLE_LIB, REPO

HEADER = os.path.join(REPO, "include", "lbdem.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lbdem_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("lbdem_create", "lbdem_init", "lbdem_dem_step", "lbdem_coupling_step", "lbdem_lb_step",
              "lbdem_fs_step", "lbdem_cycle", "lbdem_get_fields", "lbdem_total_mass"):
        assert s in syms


@pytest.mark.parametrize("path", [GPU_LIB, ORACLE_LIB], ids=["gpu-lib", "oracle-lib"])
def test_library_exports_every_symbol(path, oracle_lib):
    if path == GPU_LIB and not os.path.exists(GPU_LIB):
        pytest.skip("GPU library not built in this checkout")
    lib = ctypes.CDLL(path, mode=os.RTLD_LOCAL)
    for s in declared_symbols():
        assert hasattr(lib, s), f"{os.path.basename(path)} does not export {s}"


def test_create_rejects_bad_configuration(oracle_lib):
    p = lbdem.Params()
    p.unitLength = p.unitTime = p.unitDensity = 1.0
    p.lbSize[:] = [8.0, 8.0, 8.0]
    p.initVisc = 0.1
    p.boundary[:] = [3, 7, 7, 7, 7, 7]  # boundary code 3: neither periodic nor a wall
    p.lbSolve = 1
    p.density = p.objectDensity = 1.0
    ctx = ctypes.c_void_p()
    rc = oracle_lib.lib.lbdem_create(ctypes.byref(p), ctypes.byref(ctx))
    assert rc == 1
    assert b"boundary" in oracle_lib.lib.lbdem_last_error(ctx)
    oracle_lib.lib.lbdem_destroy(ctx)


def test_step_before_init_is_an_error(oracle_lib):
    p = lbdem.Params()
    p.unitLength = p.unitTime = p.unitDensity = 1.0
    p.lbSize[:] = [8.0, 8.0, 8.0]
    p.initVisc = 0.1
    p.boundary[:] = [7] * 6
    p.lbSolve = 1
    p.density = p.objectDensity = 1.0
    ctx = ctypes.c_void_p()
    assert oracle_lib.lib.lbdem_create(ctypes.byref(p), ctypes.byref(ctx)) == 0
    assert oracle_lib.lib.lbdem_cycle(ctx, 1) == 2
    oracle_lib.lib.lbdem_destroy(ctx)


def test_getpot_parser_and_defaults():
    text = 'lbSizeX = 24.5  # length\nrheologyModel = NEWTONIAN\nlbRestartFile = " "\n  boundary2=7\n# c = 1\n'
    v = lbdem.config_values(text)
    assert v["lbSizeX"] == 24.5 and v["rheologyModel"] == "NEWTONIAN" and v["boundary2"] == 7
    assert v["boundary0"] == 1 and v["unitLength"] == 1.0  # reference defaults
    assert "c" not in v


def test_rheology_keys(oracle_lib):
    """LB.cpp:197-245: the model name picks the keys read; BINGHAM reads no initVisc and keeps the
    default (1.0 - 0.5) / 3.0; an unknown name keeps NEWTONIAN; an unknown model code is rejected."""
    base = lbdem.config_values("initVisc = 0.3\nplasticVisc = 0.2\nyieldStress = 1e-3\nfrictionCoefFluid = 0.4\n"
                               "voellmyConstant = 2.0\nturbulenceSolve = 1\nturbConst = 0.05\n")
    for name, model in (("BINGHAM", lbdem.RHEO_BINGHAM), ("FRICTIONAL", lbdem.RHEO_FRICTIONAL),
                        ("VOELLMY", lbdem.RHEO_VOELLMY), ("NEWTONIAN", lbdem.RHEO_NEWTONIAN),
                        ("none", lbdem.RHEO_NEWTONIAN)):
        p = lbdem.params_from_values(dict(base, rheologyModel=name))
        assert p.rheologyModel == model and p.turbulenceSolve == 1 and p.turbConst == 0.05
        assert p.initVisc == ((1.0 - 0.5) / 3.0 if model == lbdem.RHEO_BINGHAM else 0.3)
        assert (p.plasticVisc, p.yieldStress) == ((0.2, 1e-3) if model == lbdem.RHEO_BINGHAM else (0.0, 0.0))
        assert p.frictionCoefFluid == (0.4 if model in (lbdem.RHEO_FRICTIONAL, lbdem.RHEO_VOELLMY) else 0.0)
        assert p.voellmyConstant == (2.0 if model == lbdem.RHEO_VOELLMY else 0.0)
    p = lbdem.Params()
    p.unitLength = p.unitTime = p.unitDensity = 1.0
    p.lbSize[:] = [8.0, 8.0, 8.0]
    p.initVisc = 0.1
    p.boundary[:] = [7] * 6
    p.lbSolve = 1
    p.density = p.objectDensity = 1.0
    p.rheologyModel = 4
    ctx = ctypes.c_void_p()
    assert oracle_lib.lib.lbdem_create(ctypes.byref(p), ctypes.byref(ctx)) == 1
    assert b"rheologyModel" in oracle_lib.lib.lbdem_last_error(ctx)
    oracle_lib.lib.lbdem_destroy(ctx)


def test_particle_and_object_readers(tmp_path):
    lbdem.write_particles(str(tmp_path / "p.dat"), [1.0, 1.2], np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]))
    ps = lbdem.read_particles(str(tmp_path / "p.dat"), scale=2.0)
    assert np.allclose(ps.r, [2.0, 2.4])
    assert np.allclose(ps.x0, np.array([[2.0, 4.0, 6.0], [8.0, 10.0, 12.0]]) - 0.05)  # DEM.cpp:101
    lbdem.write_objects(str(tmp_path / "o.dat"), [3.0], np.array([[1.0, 2.0, 3.0]]), np.array([[0.0, -0.5, 0.0]]))
    ob = lbdem.read_objects(str(tmp_path / "o.dat"))
    assert np.allclose(ob.x0, [[1.0, 2.0, 3.0]]) and np.allclose(ob.x1, [[0.0, -0.5, 0.0]])


def _box_params(n, prec):
    p = lbdem.Params()
    p.unitLength = p.unitTime = p.unitDensity = 1.0
    p.lbSize[:] = [float(n), float(n), float(n)]
    p.initVisc = 0.1
    p.boundary[:] = [4, 4, 7, 7, 4, 4]
    p.lbSolve = p.demSolve = p.freeSurface = p.forceField = 1
    p.fluidHeight = 0.75 * n
    p.density = p.objectDensity = 1.0
    p.precision = prec
    return p


def test_device_byte_estimate_is_host_side():
    """lbdem_estimate_bytes needs no device: per-node bytes follow the storage precision
    (19 populations of 8/4/2 bytes plus the per-node fields) and the estimate grows with the
    particle count; bad parameters are rejected like lbdem_create rejects them."""
    if not os.path.exists(GPU_LIB):
        pytest.skip("GPU library not built in this checkout")
    lib = lbdem.Library.load(GPU_LIB)
    ps = lbdem.ParticleSet(size=np.ones(100, np.int32), r=np.full(100, 3.0), x0=np.zeros((100, 3)))
    obs = lbdem.ObjectSet(r=np.array([8.0]), x0=np.zeros((1, 3)), x1=np.zeros((1, 3)))
    est = {prec: lbdem.estimate_bytes(lib, _box_params(96, prec), ps, obs)
           for prec in (lbdem.PREC_FP64, lbdem.PREC_FP32, lbdem.PREC_FP16)}
    N = 96 ** 3
    assert est[lbdem.PREC_FP64] > est[lbdem.PREC_FP32] > est[lbdem.PREC_FP16]
    # populations: 19 x (8 - 4) bytes per node, plus mass and 7 auxiliary fields (8 -> 4 bytes)
    assert est[lbdem.PREC_FP64] - est[lbdem.PREC_FP32] == pytest.approx(N * (19 * 4 + 8 * 4), rel=0.01)
    assert est[lbdem.PREC_FP32] - est[lbdem.PREC_FP16] == pytest.approx(N * 19 * 2, rel=0.01)
    more = lbdem.ParticleSet(size=np.ones(1000, np.int32), r=np.full(1000, 3.0), x0=np.zeros((1000, 3)))
    assert lbdem.estimate_bytes(lib, _box_params(96, lbdem.PREC_FP32), more, obs) > est[lbdem.PREC_FP32]
    bad = _box_params(96, lbdem.PREC_FP32)
    bad.boundary[0] = 3
    with pytest.raises(lbdem.LbdemError):
        lbdem.estimate_bytes(lib, bad, ps, obs)
