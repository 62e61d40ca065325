"""Batches and the run pool of the device library. A batch steps several boxes with one launch per
kernel; the pool admits independent impact runs against an HBM budget and steps them
concurrently.

Identity: every run's result in the mix equals the same run alone (its own context, lbdem_cycle
over all steps) with the same z-walk length of the fused kernel: the tests pin the walk (8 nodes)
on both sides, since the walk sets the order of the interface list and of the atomic sums, whose
round-off FP32 storage amplifies over a run, and a tuned walk depends on timings. The kernels add FP64 partial forces and the
free-surface mass surplus with atomics, whose order can differ from one execution to the next, so
"equal" is to round-off growth over the run: impactor records within 1e-9 relative (of each
record's largest entry), total fluid mass within 1e-12 relative, active node count exact."""
import numpy as np
import pytest

import lbdem
import workloads

pytestmark = pytest.mark.gpu

PRECS = (lbdem.PREC_FP32, lbdem.PREC_FP64, lbdem.PREC_FP16, lbdem.PREC_FP32, lbdem.PREC_FP32, lbdem.PREC_FP64)
STEPS = (80, 61, 40, 120, 96, 50)


@pytest.fixture(scope="module")
def runs():
    rng = np.random.default_rng(7)
    out = []
    for prec, n in zip(PRECS, STEPS):
        run = workloads.random_impact_run(rng, (32, 40, 48), n, relax=100)
        out.append((run, prec))
    return out


BATCH_TILE = 8  # z-walk the tests pin: the pool's runs, the batches' members and the runs alone


def alone(lib, run, prec, graphs=True, tile=0):
    prm = lbdem.params_from_values(run.values, prec)
    sim = lbdem.Simulation(lib, prm, run.particles, run.objects)
    sim.init()
    if tile:
        sim.set_tile(tile)
    sim.use_graphs(graphs)
    sim.cycle(run.nsteps)
    sim.synchronize()
    res = dict(objects=sim.objects(), totalMass=sim.total_mass(), nActive=sim.info()["nActive"],
               bytes=sim.device_bytes())
    sim.close()
    return res


@pytest.fixture(scope="module")
def alone_results(gpu_lib, runs):
    """The runs alone with the pinned z-walk."""
    return [alone(gpu_lib, run, prec, tile=BATCH_TILE) for run, prec in runs]


def estimate(lib, run, prec):
    return lbdem.estimate_bytes(lib, lbdem.params_from_values(run.values, prec), run.particles, run.objects)


def same(a, b, where):
    obj_a, obj_b = np.asarray(a["objects"]), np.asarray(b["objects"])
    assert obj_a.shape == obj_b.shape, where
    for k in range(obj_a.shape[0]):
        scale = max(np.abs(obj_a[k]).max(), 1e-300)
        d = np.abs(obj_a[k] - obj_b[k]).max() / scale
        print(f"{where}: impactor record rel diff {d:.3e}", flush=True)
        assert d <= 1e-9, f"{where}: impactor record differs by {d}"
    dm = abs(a["totalMass"] - b["totalMass"]) / abs(b["totalMass"])
    assert dm <= 1e-12, f"{where}: total mass differs by {dm}"
    assert a["nActive"] == b["nActive"], where


def submit_all(pool, runs, ids=None):
    out = []
    for i, (run, prec) in enumerate(runs):
        if ids is not None and i not in ids:
            continue
        out.append(pool.submit(lbdem.params_from_values(run.values, prec), run.particles, run.objects, run.nsteps))
    return out


def test_estimate_equals_allocation(gpu_lib, runs, alone_results):
    """lbdem_estimate_bytes (host-side) is exactly what lbdem_init allocates."""
    for (run, prec), res in zip(runs, alone_results):
        assert estimate(gpu_lib, run, prec) == res["bytes"], (run.edge, prec)


def test_graphs_do_not_change_results(gpu_lib, runs, alone_results):
    run, prec = runs[0]
    same(alone(gpu_lib, run, prec, graphs=False, tile=BATCH_TILE), alone_results[0], "eager vs graph replay")


@pytest.mark.parametrize("batching", [True, False], ids=["batches", "streams"])
def test_pool_results_equal_alone(gpu_lib, runs, alone_results, batching):
    """Runs of three precisions and six lengths (even and odd) submitted together: stepped in
    batches after their start-up steps (or each on its own stream), every result is the run's own."""
    budget = int(1.05 * sum(estimate(gpu_lib, r, p) for r, p in runs)) + (1 << 20) * len(runs)
    pool = lbdem.Pool(gpu_lib, budget_bytes=budget, max_concurrent=8, steps_per_chunk=16, batching=batching,
                      batch_tile=BATCH_TILE)
    ids = submit_all(pool, runs)
    pool.drain(300.0)
    st = pool.stats()
    print("pool stats", st, flush=True)
    for i, rid in enumerate(ids):
        res = pool.poll(rid)
        assert res["state"] == lbdem.RUN_DONE, res["message"]
        assert res["stepsDone"] == runs[i][0].nsteps
        assert res["nodeSteps"] == runs[i][0].nodes * runs[i][0].nsteps
        assert res["tSubmit"] <= res["tStart"] <= res["tDone"]
        assert res["tileZ"] == BATCH_TILE
        same(res, alone_results[i], f"run {i} in the mix")
    assert st["done"] == len(runs) and st["failed"] == 0
    assert st["activePeak"] >= 2, "runs did not overlap"
    assert st["bytesPeak"] <= budget and st["bytesInUse"] == 0
    if batching:
        assert st["batchBuilds"] >= 1 and st["batchedPeak"] >= 2, "no batch formed"
    else:
        assert st["batchBuilds"] == 0
    pool.close()


def test_pool_admits_by_budget(gpu_lib, runs, alone_results):
    """A budget that holds the largest run but not two of them: runs are admitted one or two at
    a time, in arrival order, and every result is still the run's own."""
    sizes = [estimate(gpu_lib, r, p) for r, p in runs]
    budget = int(1.5 * max(sizes))
    pool = lbdem.Pool(gpu_lib, budget_bytes=budget, max_concurrent=16, steps_per_chunk=32, batch_tile=BATCH_TILE)
    ids = submit_all(pool, runs)
    pool.drain(300.0)
    st = pool.stats()
    print("pool stats", st, flush=True)
    assert st["bytesPeak"] <= budget
    starts = [pool.poll(rid)["tStart"] for rid in ids]
    assert starts == sorted(starts), "admission is in arrival order"
    for i, rid in enumerate(ids):
        res = pool.poll(rid)
        assert res["state"] == lbdem.RUN_DONE, res["message"]
        same(res, alone_results[i], f"run {i} under a tight budget")
    pool.close()


def test_pool_cancel_leaves_other_runs_unchanged(gpu_lib, runs, alone_results):
    pool = lbdem.Pool(gpu_lib, budget_bytes=int(1.2 * sum(estimate(gpu_lib, r, p) for r, p in runs[:3])),
                      max_concurrent=4, steps_per_chunk=8, batch_tile=BATCH_TILE)
    ids = submit_all(pool, runs, ids={0, 1, 2})
    pool.cancel(ids[1])
    pool.drain(300.0)
    assert pool.poll(ids[1])["state"] == lbdem.RUN_CANCELLED
    for k in (0, 2):
        res = pool.poll(ids[k])
        assert res["state"] == lbdem.RUN_DONE, res["message"]
        same(res, alone_results[k], f"run {k} beside a cancelled run")
    st = pool.stats()
    assert st["cancelled"] == 1 and st["bytesInUse"] == 0
    pool.close()


def test_pool_rejects_a_run_above_the_budget(gpu_lib, runs):
    run, prec = runs[0]
    pool = lbdem.Pool(gpu_lib, budget_bytes=estimate(gpu_lib, run, prec) // 2, max_concurrent=2)
    with pytest.raises(lbdem.LbdemError, match="budget"):
        pool.submit(lbdem.params_from_values(run.values, prec), run.particles, run.objects, run.nsteps)
    pool.close()


def test_pool_destroy_with_work_outstanding(gpu_lib, runs):
    """Destroying a pool cancels queued and active runs and returns after the device is done."""
    pool = lbdem.Pool(gpu_lib, budget_bytes=int(1.2 * sum(estimate(gpu_lib, r, p) for r, p in runs)),
                      max_concurrent=2, steps_per_chunk=8)
    submit_all(pool, runs)
    pool.close()


WARM = 12  # start-up steps alone: the impactor wait, the tile tuner (8 LB steps), the first LB step


def test_batch_results_equal_alone(gpu_lib, runs, alone_results):
    """The FP32 runs (boxes of different sizes) warm up alone, advance together as one batch, and
    finish alone: every result equals the run alone."""
    idx = [i for i, (_, p) in enumerate(runs) if p == lbdem.PREC_FP32]
    sims = []
    for i in idx:
        run, prec = runs[i]
        sim = lbdem.Simulation(gpu_lib, lbdem.params_from_values(run.values, prec), run.particles, run.objects)
        sim.init()
        sim.set_tile(BATCH_TILE)
        sim.cycle(WARM)
        sims.append(sim)
    common = min(runs[i][0].nsteps for i in idx) - WARM
    common -= common % 2
    batch = lbdem.Batch(gpu_lib, sims)
    batch.cycle(common // 2 * 2 - 10)
    batch.cycle(10)
    batch.synchronize()
    batch.close()
    for k, i in enumerate(idx):
        sims[k].cycle(runs[i][0].nsteps - WARM - common)
        sims[k].synchronize()
        res = dict(objects=sims[k].objects(), totalMass=sims[k].total_mass(), nActive=sims[k].info()["nActive"])
        same(res, alone_results[i], f"run {i} through a batch of {len(idx)}")
        sims[k].close()


def test_batch_rejects_members_that_cannot_share_launches(gpu_lib, runs):
    run32, run64 = runs[0][0], runs[1][0]
    a = lbdem.Simulation(gpu_lib, lbdem.params_from_values(run32.values, lbdem.PREC_FP32), run32.particles, run32.objects)
    b = lbdem.Simulation(gpu_lib, lbdem.params_from_values(run64.values, lbdem.PREC_FP64), run64.particles, run64.objects)
    a.init()
    b.init()
    with pytest.raises(lbdem.LbdemError):  # start-up steps not done
        lbdem.Batch(gpu_lib, [a, b])
    a.cycle(WARM)
    b.cycle(WARM)
    with pytest.raises(lbdem.LbdemError):  # precisions differ
        lbdem.Batch(gpu_lib, [a, b])
    b.cycle(1)
    with pytest.raises(lbdem.LbdemError):  # step parities differ
        lbdem.Batch(gpu_lib, [a, a])
    a.close()
    b.close()


def test_batch_of_mixed_sizes_equals_alone(gpu_lib):
    """Boxes whose fused-kernel grids differ by 8x (48^3: 72 blocks, 96^3: 576) share each launch,
    whose grid is the largest box's (the surplus blocks of the small boxes return at once); every
    result still equals the run alone (same z-walk)."""
    rng = np.random.default_rng(11)
    runs = [workloads.random_impact_run(rng, (e, e, e), 60, relax=100) for e in (48, 96, 48)]
    prec = lbdem.PREC_FP32
    ref = [alone(gpu_lib, run, prec, tile=BATCH_TILE) for run in runs]
    sims = []
    for run in runs:
        sim = lbdem.Simulation(gpu_lib, lbdem.params_from_values(run.values, prec), run.particles, run.objects)
        sim.init()
        sim.set_tile(BATCH_TILE)
        sim.cycle(WARM)
        sims.append(sim)
    batch = lbdem.Batch(gpu_lib, sims)
    batch.cycle(60 - WARM)
    batch.synchronize()
    batch.close()
    for k, sim in enumerate(sims):
        res = dict(objects=sim.objects(), totalMass=sim.total_mass(), nActive=sim.info()["nActive"])
        same(res, ref[k], f"box {runs[k].edge} through a mixed-size batch")
        sim.close()


def test_batch_tile_tuner(gpu_lib):
    """lbdem_batch_set_tile(batch, 0) times each candidate walk on one step pair of the batch and
    keeps the fastest; the runs continue correctly (results equal the runs alone at that walk)."""
    rng = np.random.default_rng(5)
    runs = [workloads.random_impact_run(rng, (e, e, e), 60, relax=100) for e in (48, 64)]
    prec = lbdem.PREC_FP32
    sims = []
    for run in runs:
        sim = lbdem.Simulation(gpu_lib, lbdem.params_from_values(run.values, prec), run.particles, run.objects)
        sim.init()
        sim.set_tile(BATCH_TILE)
        sim.cycle(WARM)
        sims.append(sim)
    batch = lbdem.Batch(gpu_lib, sims)
    batch.set_tile(0)
    assert not batch.tile()["tuned"]
    batch.cycle(6)  # three candidates, one pair each
    t = batch.tile()
    print("batch tile", t, flush=True)
    assert t["tuned"] and t["tileZ"] in (2, 4, 8)
    batch.cycle(60 - WARM - 6)
    batch.synchronize()
    batch.close()
    for run, sim in zip(runs, sims):
        assert sim.info()["lbTime"] == 60
        assert np.isfinite(sim.total_mass())
        sim.close()


def test_batch_kernel_timing(gpu_lib):
    """lbdem_batch_profile_enable / _read: with timing on, every launch of the mixed-size batch is
    timed (the fused kernel once per step for all three boxes) and the runs still equal the runs
    alone."""
    rng = np.random.default_rng(11)
    runs = [workloads.random_impact_run(rng, (e, e, e), 60, relax=100) for e in (48, 96, 48)]
    prec = lbdem.PREC_FP32
    ref = [alone(gpu_lib, run, prec, tile=BATCH_TILE) for run in runs]
    sims = []
    for run in runs:
        sim = lbdem.Simulation(gpu_lib, lbdem.params_from_values(run.values, prec), run.particles, run.objects)
        sim.init()
        sim.set_tile(BATCH_TILE)
        sim.cycle(WARM)
        sims.append(sim)
    batch = lbdem.Batch(gpu_lib, sims)
    batch.cycle(10)
    batch.profile(True)
    assert batch.profile_read() == []
    batch.cycle(20)
    rows = batch.profile_read()
    print("batch kernels", [(r["kernel"], r["launches"]) for r in rows[:6]], flush=True)
    fused = [r for r in rows if r["kernel"].startswith("k_fused<float")]
    assert fused and sum(r["launches"] for r in fused) == 20
    assert all(r["ms"] > 0 and r["launches"] > 0 for r in rows)
    assert len({r["kernel"] for r in rows}) == len(rows)
    batch.profile(False)
    batch.cycle(60 - WARM - 30)
    batch.synchronize()
    batch.close()
    for k, sim in enumerate(sims):
        res = dict(objects=sim.objects(), totalMass=sim.total_mass(), nActive=sim.info()["nActive"])
        same(res, ref[k], f"box {runs[k].edge} through a timed batch")
        sim.close()
