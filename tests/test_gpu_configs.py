"""The benchmark configurations on the device, each checked against something independent of the
run itself:
  * the reference-scale run (64^3, 480 spheres of radius 3, impactor r 8): FP64 against the oracle,
    step by step, and FP32 offsets within their tolerance;
  * the headline run at full size (256 x 320 x 256, 31 000 spheres of radius 4, impactor r 16): FP64
    against the oracle over 30 steps, and the Verlet list of the first rebuild against
    a brute-force search (scipy's periodic k-d tree, then the reference's minimum-image test);
  * the precision sweep on the headline run: FP32 offsets and FP16 storage against FP64 over the
    5000 steps of the run (impactor depth of objPosition.dat, fluid mass drift);
  * the batched workload (64 boxes of edge 48-128, 2000 steps, FP32): every box in the batch
    against the same box alone;
  * a pool batch holding an oracle case (FP64) beside other runs: the member against the oracle.
Tolerances are those of test_gpu_parity (FP64: populations 1e-13 absolute, records 1e-10 relative;
FP32: 2e-7 / 1e-4) unless stated at the test."""
import numpy as np
import pytest

import cases
import lbdem
import workloads
from test_gpu_parity import TOL, compare, rel

pytestmark = pytest.mark.gpu


def _pair(oracle_lib, gpu_lib, vals, ps, obs, prec):
    a = lbdem.Simulation(oracle_lib, lbdem.params_from_values(vals, lbdem.PREC_FP64), ps, obs)
    a.init()
    b = lbdem.Simulation(gpu_lib, lbdem.params_from_values(vals, prec), ps, obs)
    b.init()
    return a, b


@pytest.mark.parametrize("prec", [lbdem.PREC_FP64, lbdem.PREC_FP32])
def test_reference_scale_run_matches_oracle(oracle_lib, gpu_lib, prec):
    vals, ps, obs = workloads.impact_box(**workloads.CONFIG0)
    a, b = _pair(oracle_lib, gpu_lib, vals, ps, obs, prec)
    print(f"64^3 run, {len(ps.r)} spheres, prec {prec}", flush=True)
    compare(cases.state(a), cases.state(b), prec, "init")
    done = 0
    for target in (1, 10, 40, 120):
        a.cycle(target - done)
        b.cycle(target - done)
        b.synchronize()
        done = target
        compare(cases.state(a), cases.state(b), prec, f"step {done}")
    print("neighbour-list builds", a.info()["neighborListBuilds"], b.info()["neighborListBuilds"], flush=True)
    a.close()
    b.close()


def _brute_verlet(x, r, L):
    """Partner sets of the reference's criterion (minimum image in every axis of the DEM box,
    |d|^2 <= (r1 + r2 + 2)^2, DEM.cpp:521-531) from candidate pairs of a periodic k-d tree."""
    from scipy.spatial import cKDTree
    L = np.asarray(L, dtype=np.float64)
    xw = np.mod(x, L)
    cut = 2.0 * float(r.max()) + 2.0
    pairs = cKDTree(xw, boxsize=L).query_pairs(cut + 1e-6, output_type="ndarray")
    d = x[pairs[:, 0]] - x[pairs[:, 1]]
    d = d - cases._n_int(d / L) * L
    r12 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    c = r[pairs[:, 0]] + r[pairs[:, 1]] + 2.0
    keep = pairs[r12 <= c * c]
    out = [[] for _ in range(len(r))]
    for i, j in keep:
        out[i].append(int(j))
        out[j].append(int(i))
    return [sorted(p) for p in out]


def test_headline_full_size_fp64_matches_oracle(oracle_lib, gpu_lib):
    vals, ps, obs = workloads.impact_box(**workloads.HEADLINE)
    a, b = _pair(oracle_lib, gpu_lib, vals, ps, obs, lbdem.PREC_FP64)
    print(f"headline run at full size: {b.info()['size']}, {len(ps.r)} spheres", flush=True)
    L = cases.dem_size(vals)
    done = 0
    for target in (1, 30):
        a.cycle(target - done)
        b.cycle(target - done)
        b.synchronize()
        done = target
        A, B = cases.state(a), cases.state(b)
        compare(A, B, lbdem.PREC_FP64, f"step {done}")
        del A, B
    # the Verlet list of the next rebuild against brute force from the positions it is built from
    # (those at the start of its step); the spheres start at rest, so the first rebuild after
    # initialisation comes after a few tens of steps
    for _ in range(120):
        x_before = b.particles()[:, 0:3].copy()
        n0 = b.info()["neighborListBuilds"]
        b.cycle(1)
        b.synchronize()
        if b.info()["neighborListBuilds"] > n0:
            break
    else:
        pytest.fail("no neighbour-list rebuild within 120 steps")
    cnt, lst = b.neighbors()
    expect = _brute_verlet(x_before, ps.r, L)
    bad = [p for p in range(len(cnt)) if lst[p, :cnt[p]].tolist() != expect[p]]
    assert not bad, f"{len(bad)} lists differ, first {bad[:5]}"
    a.close()
    b.close()


def test_precision_sweep_depth_and_mass(gpu_lib):
    """The headline run in FP64, FP32 offsets and FP16 storage over its 5000 steps: impactor depth
    (fluidHeight + r - y, IO.cpp:1323) of FP32 within 0.1 % of FP64's and FP16's within 1 % at the
    end and at the deepest point, and each run's fluid-mass drift within 1e-3 of FP64's."""
    vals, ps, obs = workloads.impact_box(**workloads.HEADLINE)
    fh, r_imp = float(vals["fluidHeight"]), float(obs.r[0])
    sims = {}
    for prec in (lbdem.PREC_FP64, lbdem.PREC_FP32, lbdem.PREC_FP16):
        s = lbdem.Simulation(gpu_lib, lbdem.params_from_values(vals, prec), ps, obs)
        s.init()
        s.cycle(1)  # mass baseline after the first coupling step (it takes the particle nodes out)
        sims[prec] = (s, s.total_mass(), [])
    done = 1
    for step in range(500, 5001, 500):
        for prec, (s, m0, depth) in sims.items():
            s.cycle(step - done)
        done = step
        for prec, (s, m0, depth) in sims.items():
            depth.append(fh + r_imp - float(s.objects()[0, lbdem.O["X0"] + 1]))
        print(f"step {step}: depth " + " ".join(f"{sims[p][2][-1]:.5f}" for p in sims), flush=True)
    d64 = np.array(sims[lbdem.PREC_FP64][2])
    k = int(np.argmax(np.abs(d64)))
    drift = {p: (s.total_mass() - m0) / m0 for p, (s, m0, _) in sims.items()}
    for prec, lim in ((lbdem.PREC_FP32, 1e-3), (lbdem.PREC_FP16, 1e-2)):
        d = np.array(sims[prec][2])
        assert abs(d[-1] - d64[-1]) <= lim * abs(d64[-1]), (prec, d[-1], d64[-1])
        assert abs(d[k] - d64[k]) <= lim * abs(d64[k]), (prec, d[k], d64[k])
        assert abs(drift[prec] - drift[lbdem.PREC_FP64]) <= 1e-3, drift
    assert d64[-1] > 5.0  # the impactor is well inside the bed by the end
    for s, _, _ in sims.values():
        s.close()


def test_batched_workload_equals_alone(gpu_lib):
    """64 boxes of edge 48-128, 2000 steps each, FP32: each box stepped in one batch with the others
    equals the box alone, both with the 8-node walk of the fused kernel. FP32 storage amplifies the
    round-off of atomic sums whose order differs between executions over 2000 steps until a
    free-surface conversion, an impactor cover or uncover (which the reference does not book, LB.cpp:
    909-938) or a particle flag goes the other way, each moving about one node's mass (rho0 = 1),
    so: impactor position and velocity within 1e-6 of the record's largest entry, total fluid mass
    within 8 nodes' mass, active node count within 0.1 %."""
    runs = workloads.batched_runs(64, 2000)
    prec = lbdem.PREC_FP32
    sims = []
    for run in runs:
        s = lbdem.Simulation(gpu_lib, lbdem.params_from_values(run.values, prec), run.particles, run.objects)
        s.init()
        s.set_tile(8)
        s.cycle(4)  # start-up steps alone
        sims.append(s)
    for s in sims:
        s.synchronize()
    batch = lbdem.Batch(gpu_lib, sims)
    batch.set_tile(8)
    batch.cycle(1996)
    batch.synchronize()
    res_b = [(s.objects()[0, :6].copy(), s.total_mass(), s.info()["nActive"]) for s in sims]
    batch.close()
    for s in sims:
        s.close()
    print(f"batch of {len(runs)} done", flush=True)
    for i, run in enumerate(runs):
        s = lbdem.Simulation(gpu_lib, lbdem.params_from_values(run.values, prec), run.particles, run.objects)
        s.init()
        s.set_tile(8)
        s.cycle(2000)
        s.synchronize()
        ob, mb, nb = res_b[i]
        oa = s.objects()[0, :6]
        assert np.abs(oa - ob).max() <= 1e-6 * np.abs(oa).max(), (i, run.edge, oa, ob)
        assert abs(s.total_mass() - mb) <= 8.0, (i, s.total_mass(), mb)
        assert abs(s.info()["nActive"] - nb) <= 1e-3 * nb, (i, s.info()["nActive"], nb)
        s.close()


def test_batch_member_matches_oracle(oracle_lib, gpu_lib, case_root):
    """An oracle case (the suspension, FP64) stepped in a batch beside two other runs: its full
    state against the oracle's after the batch's steps."""
    d = cases.make_case(case_root, "suspension")
    a = cases.open_sim(oracle_lib, d)
    m = cases.open_sim(gpu_lib, d, precision=lbdem.PREC_FP64)
    others = []
    for name in ("periodic", "moving"):  # runs with spheres: the same step phase after the same steps
        others.append(cases.open_sim(gpu_lib, cases.make_case(case_root, name), precision=lbdem.PREC_FP64))
    for s in [m] + others:
        s.set_tile(4)
        s.cycle(6)
        s.synchronize()
    a.cycle(6)
    batch = lbdem.Batch(gpu_lib, [m] + others)
    done = 6
    for target in (16, 46):
        batch.cycle(target - done)
        batch.synchronize()
        a.cycle(target - done)
        done = target
        compare(cases.state(a), cases.state(m), lbdem.PREC_FP64, f"batch member, step {done}")
    batch.close()
    for s in [a, m] + others:
        s.close()
