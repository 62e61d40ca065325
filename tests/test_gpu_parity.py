"""MI355X library against the FP64 oracle on identical inputs, step by step, through the C ABI.

Tolerances (absolute unless marked relative; populations are O(1e-1), masses O(1)):
  FP64 storage  : node types / particle flags / solid indices / Verlet lists identical;
                  populations 1e-13, mass and density 1e-11, particle / impactor records 1e-10 relative
  FP32 offsets  : flags identical over the window; populations 2e-7, mass 5e-6,
                  particle records 1e-4 relative, impactor 1e-6 relative
  FP16 storage  : flags identical over the window; populations 5e-5, mass 2e-4,
                  particle records 1e-3 relative
  Stress diagnostics (each of the five global tensors, STotal, relative to their largest entry):
                  FP64 1e-9, FP32 2e-2; disp 1e-10 (FP64) and 1e-4 (FP32) absolute
The reference's own golden fixtures are also compared with the FP64 library directly."""
import json
import os

import numpy as np
import pytest

import cases
import lbdem

pytestmark = pytest.mark.gpu

TOL = {
    lbdem.PREC_FP64: dict(f=1e-13, mass=1e-11, n=1e-11, part=1e-10, obj=1e-10, wall=1e-8),
    lbdem.PREC_FP32: dict(f=2e-7, mass=5e-6, n=5e-6, part=1e-4, obj=1e-6, wall=1e-3),
    lbdem.PREC_FP16: dict(f=5e-5, mass=2e-4, n=2e-4, part=1e-3, obj=1e-4, wall=1e-2),
}


def rel(a, b):
    if a.size == 0:
        return 0.0
    return float(np.abs(a - b).max() / max(np.abs(a).max(), 1e-300))


def compare(A, B, prec, where, free_surface=True):
    t = TOL[prec]
    for k in ("type", "pflag", "solid"):
        bad = np.nonzero(A[k] != B[k])[0]
        assert bad.size == 0, f"{where}: {k} differs at {bad[:8]}"
    act = (A["type"] == 0) | (A["type"] == 3)
    df = np.abs(A["f"][:, act] - B["f"][:, act]).max()
    assert df <= t["f"], f"{where}: populations differ by {df}"
    dm = np.abs(A["mass"][act] - B["mass"][act]).max()
    assert dm <= t["mass"], f"{where}: mass differs by {dm}"
    if free_surface:
        dn = np.abs(A["n"][act] - B["n"][act]).max()
        assert dn <= t["n"], f"{where}: density differs by {dn}"
    assert rel(A["particles"], B["particles"]) <= t["part"], f"{where}: particle records"
    assert rel(A["objects"], B["objects"]) <= t["obj"], f"{where}: impactor records"
    assert rel(A["walls"], B["walls"]) <= t["wall"], f"{where}: wall records"
    assert rel(A["bodies"], B["bodies"]) <= t["obj"], f"{where}: body records"
    assert np.array_equal(A["ncount"], B["ncount"]), f"{where}: neighbour counts"
    # partner multisets: the reference (and the oracle) list partners in cell-walk order, the
    # library in ascending order
    for p in range(len(A["ncount"])):
        c = A["ncount"][p]
        assert np.array_equal(np.sort(A["nlist"][p, :c]), B["nlist"][p, :c]), f"{where}: neighbours of {p}"


def run_pair(oracle_lib, gpu_lib, case_root, name, prec, steps, every=1):
    d = cases.make_case(case_root, name)
    a = cases.open_sim(oracle_lib, d)
    b = cases.open_sim(gpu_lib, d, precision=prec)
    fs = int(cases._case_values(name)[0]["freeSurfaceSolve"]) != 0
    print(f"case={name} prec={prec} steps={steps}", flush=True)
    compare(cases.state(a), cases.state(b), prec, f"{name} init", fs)
    done = 0
    while done < steps:
        n = min(every, steps - done)
        a.cycle(n)
        b.cycle(n)
        b.synchronize()
        done += n
        compare(cases.state(a), cases.state(b), prec, f"{name} step {done}", fs)
    a.close()
    b.close()


@pytest.mark.parametrize("name", cases.CASES)
def test_fp64_matches_oracle(oracle_lib, gpu_lib, case_root, name):
    run_pair(oracle_lib, gpu_lib, case_root, name, lbdem.PREC_FP64, 40)


@pytest.mark.parametrize("name", cases.CELL_CASES)
def test_fp64_cell_branch_matches_oracle(oracle_lib, gpu_lib, case_root, name):
    """The DEM link-cell neighbour list on the device (counting sort into cells, 27-cell gather):
    "cells" has 6 cells per axis and spheres that cross the x and z seams at step 25 (60 steps,
    compared every 4); thin_z2 / thin_z1 / thin_z0 have 2, 1 and 0 cells along z (the reference
    lists pairs once per visit of a repeated z-cell, and builds no list without z-cells)."""
    steps, every = (60, 4) if name == "cells" else (40, 5)
    run_pair(oracle_lib, gpu_lib, case_root, name, lbdem.PREC_FP64, steps, every=every)


def test_fp32_cell_branch_within_tolerance(oracle_lib, gpu_lib, case_root):
    run_pair(oracle_lib, gpu_lib, case_root, "cells", lbdem.PREC_FP32, 40, every=10)


def test_fp64_long_window(oracle_lib, gpu_lib, case_root):
    run_pair(oracle_lib, gpu_lib, case_root, "suspension", lbdem.PREC_FP64, 300, every=25)


@pytest.mark.parametrize("name", ["fs_impactor", "periodic", "moving", "startup", "two_impactors"])
def test_graph_replays_match_oracle(oracle_lib, gpu_lib, case_root, name):
    """cycle(n) with n >= 2 replays a captured pair of steady-state steps (HIP graph): 10-step
    chunks with downloads and synchronisation between replays, every chunk against the oracle."""
    run_pair(oracle_lib, gpu_lib, case_root, name, lbdem.PREC_FP64, 60, every=10)


@pytest.mark.parametrize("tile_z", [1, 2, 4, 8])
@pytest.mark.parametrize("name", ["suspension", "periodic"])
def test_every_tile_matches_oracle(oracle_lib, gpu_lib, case_root, name, tile_z):
    """The fused kernel with each z-walk length the tile tuner can pick."""
    d = cases.make_case(case_root, name)
    a = cases.open_sim(oracle_lib, d)
    b = cases.open_sim(gpu_lib, d, precision=lbdem.PREC_FP64)
    b.set_tile(tile_z)
    print(f"case={name} tile_z={tile_z}", flush=True)
    for step in range(3):
        a.cycle(10)
        b.cycle(10)
        b.synchronize()
        compare(cases.state(a), cases.state(b), lbdem.PREC_FP64, f"{name} tile {tile_z} step {10 * (step + 1)}")
    assert b.tuning()["tile"] == [64, 4, tile_z]
    a.close()
    b.close()


def test_tile_tuner_picks_a_candidate(gpu_lib, case_root):
    d = cases.make_case(case_root, "suspension")
    sim = cases.open_sim(gpu_lib, d, precision=lbdem.PREC_FP32)
    assert not sim.tuning()["tuned"]
    sim.cycle(12)
    sim.synchronize()
    t = sim.tuning()
    print("tuning", t, flush=True)
    assert t["tuned"] and t["tile"][2] in (1, 2, 4, 8)
    assert all(ms > 0 for ms in t["candidates"].values())
    assert t["candidates"][t["tile"][2]] == min(t["candidates"].values())
    sim.close()


@pytest.mark.parametrize("name", ["suspension", "units", "periodic"])
def test_fp32_offsets_within_tolerance(oracle_lib, gpu_lib, case_root, name):
    run_pair(oracle_lib, gpu_lib, case_root, name, lbdem.PREC_FP32, 40, every=5)


def test_fp16_storage_within_tolerance(oracle_lib, gpu_lib, case_root):
    run_pair(oracle_lib, gpu_lib, case_root, "suspension", lbdem.PREC_FP16, 30, every=5)


@pytest.mark.parametrize("prec", [lbdem.PREC_FP64, lbdem.PREC_FP32])
@pytest.mark.parametrize("name", sorted(cases.RHEOLOGY_CASES))
def test_rheology_matches_oracle(oracle_lib, gpu_lib, case_root, name, prec):
    """The fused kernel's variant with per-node viscosity: BINGHAM with the Smagorinsky term,
    FRICTIONAL, VOELLMY and Newtonian with the Smagorinsky term, 40 steps, every 5 against the oracle."""
    run_pair(oracle_lib, gpu_lib, case_root, name, prec, 40, every=5)


@pytest.mark.parametrize("prec", [lbdem.PREC_FP64, lbdem.PREC_FP32])
@pytest.mark.parametrize("name", sorted(cases.BODY_CASES))
def test_bodies_match_oracle(oracle_lib, gpu_lib, case_root, name, prec):
    """Spring-connected bodies (DEM.cpp:798-869, 2730-2740): WCA and cubic springs with dashpots
    pulling the impactor, 60 steps compared every 10 (body and impactor records included)."""
    run_pair(oracle_lib, gpu_lib, case_root, name, prec, 60, every=10)


@pytest.mark.parametrize("prec", [lbdem.PREC_FP64, lbdem.PREC_FP32])
@pytest.mark.parametrize("name", ["dyn_walls", "wall_codes", "slip_walls", "slip_dyn_walls"])
def test_wall_codes_match_oracle(oracle_lib, gpu_lib, case_root, name, prec):
    """Boundary codes 5-8 (every wall of boundary0's kind, DEM.cpp:430-446): all faces moving walls
    (type 8, zero velocity); a slip floor with a moving lid under a periodic boundary0 (both static
    walls); slip walls on x and y (the slip pass after the fused kernel, LB.cpp:1454-1482); 40 steps,
    every 10."""
    run_pair(oracle_lib, gpu_lib, case_root, name, prec, 40, every=10)


def test_restart_viscosities_match_oracle(oracle_lib, gpu_lib, case_root):
    """A Newtonian restart whose records carry viscosities other than the run's: every node relaxes
    with its own (node.cpp:40-50, 181-207), new interface nodes take their donor's (LB.cpp:1719) and
    uncovered impactor nodes the run's (LB.cpp:928-934); 30 steps against the oracle, and the
    viscosities read back."""
    d = cases.make_case(case_root, "restart")
    vals, prm, ps, obs = lbdem.load_case(d)
    size, types, rec = lbdem.read_recycle(os.path.join(d, cases.RESTART_FILE))
    rec = rec.copy()
    rec[:, 5] *= 1.0 + 0.5 * (np.arange(len(rec)) % 3)
    sims = []
    for lib in (oracle_lib, gpu_lib):
        _, prm, ps, obs = lbdem.load_case(d, precision=lbdem.PREC_FP64)
        sim = lbdem.Simulation(lib, prm, ps, obs)
        sim.set_restart(size, types, rec)
        sim.init()
        sims.append(sim)
    a, b = sims
    print("restart with per-node viscosities", flush=True)
    va, vb = a.velocity()[1], b.velocity()[1]
    assert np.array_equal(va, vb) and len(np.unique(va[va > 0])) == 3
    for step in range(6):
        a.cycle(5)
        b.cycle(5)
        b.synchronize()
        compare(cases.state(a), cases.state(b), lbdem.PREC_FP64, f"restart with viscosities, step {5 * (step + 1)}")
    for sim in (a, b):
        sim.store_velocity()
        sim.cycle(1)
    va, vb = a.velocity()[1], b.velocity()[1]
    assert np.array_equal(va, vb)
    a.close()
    b.close()


STRESS_TOL = {lbdem.PREC_FP64: (1e-9, 1e-10), lbdem.PREC_FP32: (2e-2, 1e-4)}
STRESS_NAMES = ("totalStress", "hydroStress", "lubStress", "contactStress", "electroStress")


def compare_particle_stress(A, B, prec, where):
    """n x 44 records of lbdem_get_particle_stress (A the oracle or the reference): each tensor
    relative to its largest entry, overlap and disp absolutely, the counts exactly (FP64)."""
    tol, dtol = STRESS_TOL[prec]
    A, B = np.asarray(A), np.asarray(B)
    assert A.shape == B.shape, f"{where}: particle stress shape"
    if A.size == 0:
        return
    for k, nm in enumerate(("STotal", "SElas", "SVisc", "SVisc_h")):
        assert rel(A[:, 9 * k:9 * k + 9], B[:, 9 * k:9 * k + 9]) <= tol, f"{where}: {nm}"
    assert np.abs(A[:, 36:42] - B[:, 36:42]).max() <= max(dtol, tol * np.abs(A[:, 36:39]).max()), f"{where}: overlap, disp"
    if prec == lbdem.PREC_FP64:
        assert np.array_equal(A[:, 42:44], B[:, 42:44]), f"{where}: contact and lubrication counts"


def compare_stress(A, B, prec, where):
    """A, B: (global 5x3x3, STotal n x 3 x 3, disp n x 3), A the oracle or the reference."""
    tol, dtol = STRESS_TOL[prec]
    for k, nm in enumerate(STRESS_NAMES):
        assert rel(np.asarray(A[0][k]), np.asarray(B[0][k])) <= tol, f"{where}: {nm} {A[0][k]} vs {B[0][k]}"
    assert rel(np.asarray(A[1]), np.asarray(B[1])) <= tol, f"{where}: STotal"
    if np.asarray(A[2]).size:
        assert np.abs(np.asarray(A[2]) - np.asarray(B[2])).max() <= dtol, f"{where}: disp"


@pytest.mark.parametrize("prec", [lbdem.PREC_FP64, lbdem.PREC_FP32])
@pytest.mark.parametrize("name", ["suspension", "cells", "startup", "units", "body_cubic", "periodic"])
def test_stress_diagnostics_match_oracle(oracle_lib, gpu_lib, case_root, name, prec):
    """SHydro (the stresslet pass after the fused kernel), the pair, wall and impactor stresses of
    the force kernel and the displacement of the integration, against the oracle every 10 steps."""
    d = cases.make_case(case_root, name)
    a = cases.open_sim(oracle_lib, d, stress=True)
    b = cases.open_sim(gpu_lib, d, precision=prec, stress=True)
    print(f"case={name} prec={prec} stress", flush=True)
    for step in range(10, 41, 10):
        a.cycle(10)
        b.cycle(10)
        b.synchronize()
        A, B = a.stress(), b.stress()
        assert np.abs(A[0][1]).max() > 0 or name == "periodic", "the window has hydrodynamic stress"
        compare_stress(A, B, prec, f"{name} step {step}")
        compare_particle_stress(a.particle_stress(), b.particle_stress(), prec, f"{name} step {step}")
    a.close()
    b.close()


def test_stress_needs_enabling(gpu_lib, case_root):
    sim = cases.open_sim(gpu_lib, cases.make_case(case_root, "suspension"))
    with pytest.raises(lbdem.LbdemError):
        sim.stress()
    with pytest.raises(lbdem.LbdemError):
        sim.set_stress()
    sim.close()
    # a context with the stress diagnostics cannot join a batch
    sim = cases.open_sim(gpu_lib, cases.make_case(case_root, "suspension"), stress=True)
    with pytest.raises(lbdem.LbdemError, match="stress"):
        lbdem.Batch(gpu_lib, [sim])
    sim.close()


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["suspension", "fs_impactor", "periodic", "units", "moving", "startup", "two_impactors",
                                  "cells", "thin_z2", "thin_z1", "thin_z0", "restart", "bingham", "frictional",
                                  "voellmy", "body_wca", "body_cubic", "dyn_walls", "wall_codes", "slip_walls",
                                  "slip_dyn_walls"])
def test_fp64_matches_reference_fixtures(gpu_lib, case_root, name):
    with open(os.path.join(GOLD, "manifest.json")) as fh:
        steps = sorted(m["step"] for m in json.load(fh).values() if m["case"] == name)
    d = cases.make_case(case_root, name)
    with np.load(os.path.join(GOLD, f"{name}_s{steps[0]}.npz"), allow_pickle=False) as g0:
        stress = "stress" in g0.files
    sim = cases.open_sim(gpu_lib, d, precision=lbdem.PREC_FP64, stress=stress)
    done = 0
    for step in steps:
        sim.cycle(step - done)
        done = step
        g = np.load(os.path.join(GOLD, f"{name}_s{step}.npz"), allow_pickle=False)
        st = cases.state(sim)
        if stress:
            n = g["particles"].shape[0]
            compare_stress((g["stress"].reshape(5, 3, 3), g["stotal"].reshape(n, 3, 3), g["disp"]), sim.stress(),
                           lbdem.PREC_FP64, f"{name} step {step} (reference)")
            if "pstress" in g:
                compare_particle_stress(g["pstress"], sim.particle_stress(), lbdem.PREC_FP64,
                                        f"{name} step {step} (reference)")
        assert np.array_equal(st["type"], g["type"]) and np.array_equal(st["pflag"], g["pflag"])
        assert np.array_equal(st["solid"], g["solid"])
        act = g["active"]
        assert np.abs(st["f"][:, act] - g["f_active"]).max() <= 1e-13
        assert np.abs(st["mass"][act] - g["mass_active"]).max() <= 1e-11
        assert rel(g["particles"], st["particles"]) <= 1e-10
        assert rel(g["objects"], st["objects"]) <= 1e-10
        if "bodies" in g:
            assert rel(g["bodies"], st["bodies"]) <= 1e-10
        assert abs(st["totalMass"] - float(g["totalMass"])) <= 1e-11 * abs(float(g["totalMass"]))
        assert np.array_equal(st["ncount"], g["ncount"])
        for p in range(len(g["ncount"])):
            c = g["ncount"][p]
            assert np.array_equal(st["nlist"][p, :c], np.sort(g["nlist"][p, :c])), f"{name} step {step}: list of {p}"
    sim.close()


def test_node_velocity_matches_oracle(oracle_lib, gpu_lib, case_root):
    """lbdem_store_velocity / lbdem_get_velocity (what IO reads as node->u and node->visc): after
    initialisation, after a restart and after a stored LB step, every active node's velocity and
    viscosity equal the oracle's (FP64); an LB step without the store leaves none to read."""
    for name, warm in (("suspension", 12), ("restart", 5), ("voellmy", 12)):
        d = cases.make_case(case_root, name)
        a = cases.open_sim(oracle_lib, d)
        b = cases.open_sim(gpu_lib, d, precision=lbdem.PREC_FP64)
        print(f"case={name}", flush=True)
        ua, va = a.velocity()
        ub, vb = b.velocity()
        assert np.abs(ua - ub).max() <= 1e-15 and np.array_equal(va, vb), f"{name}: initial velocity"
        a.cycle(warm)
        b.cycle(warm)
        for sim in (a, b):
            sim.store_velocity()
            sim.cycle(1)
        ua, va = a.velocity()
        ub, vb = b.velocity()
        scale = max(np.abs(ua).max(), 1e-30)
        assert np.abs(ua - ub).max() <= 1e-12 * scale, f"{name}: stored velocity"
        assert np.abs(va - vb).max() <= 1e-12 * va.max() and np.array_equal(va > 0, vb > 0), f"{name}: viscosity"
        b.cycle(1)
        with pytest.raises(lbdem.LbdemError):
            b.velocity()
        a.close()
        b.close()


def test_device_errors_are_reported(gpu_lib, case_root):
    """An unsupported configuration fails at create time with a message, before any launch."""
    d = cases.make_case(case_root, "suspension")
    vals, prm, ps, obs = lbdem.load_case(d)
    prm.boundary[0] = 3
    with pytest.raises(lbdem.LbdemError, match="boundary"):
        lbdem.Simulation(gpu_lib, prm, ps, obs)
