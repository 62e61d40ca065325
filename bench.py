#!/usr/bin/env python3
"""Benchmark of the coupled LB/DEM step on one MI355X.

Workload (N=1): the headline impact run — a 256 x 256 x 320 box (x, z horizontal and periodic,
y vertical with floor and lid walls), fluid to height 256 with a free surface, 31 000 spheres of
radius 4 (phi ~ 0.5, seeded packing), an impactor of radius 16 at height 290 falling at 0.01
lattice units per step, tau = 0.8. A step is one full coupled step (DEM, coupling, fused LB collide-stream,
free surface). Inputs are resident on the device when the timed region starts.

Prints one JSON line: MLUPS over active (fluid + interface) cells for the FP32-offset step, the
FP64 step on the same state for comparison, the fused kernel's HBM roofline fraction and a CPU
baseline (the reference program compiled from its sources, on a bounded sample); then the
batched workload (64 boxes of edge 48-128 submitted together to the run pool) and the served
workload (Poisson arrivals of runs with edge 64-192 and 1000-8000 steps at 0.8 of the batched
capacity: p50/p99 latency and runs/hour, beside the same requests served one at a time with
plain launches).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "lbm-dem_amd", "python"))

import numpy as np  # noqa: E402

import lbdem  # noqa: E402
import workloads  # noqa: E402

PEAK_HBM_GBS = 8000.0
# HBM bytes per fused-kernel launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs on this
# workload (scaled by the same counters on a plain copy of known size)
TRAFFIC_FILE = os.path.join(HERE, "profiles", "k_fused_impact256_fp32_hbm_traffic.json")
LBDEM_E_TYPE = 5  # include/lbdem.h: streaming met a neighbour type it cannot handle
BYTES_PER_CELL = {lbdem.PREC_FP64: 38 * 8 + 1, lbdem.PREC_FP32: 38 * 4 + 1, lbdem.PREC_FP16: 38 * 2 + 1}
PREC_NAME = {lbdem.PREC_FP64: "fp64", lbdem.PREC_FP32: "fp32", lbdem.PREC_FP16: "fp16"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(vals, ps, obs, prec, steps, warmup, run_steps=5000, timed_from=1500, every=100):
    """One headline run of ``run_steps`` coupled steps in precision ``prec``.

    Untimed before the timed window: the ``warmup`` steps asked for, then step pairs until the
    context's start-up is over (the fused kernel's tile tuner has picked its walk and the steady
    step pair is captured as a HIP graph), then steps up to ``timed_from`` (the impactor meets the
    surface between steps 1000 and 1500, so the window lies in the impact phase: splash
    conversions, impactor uncover/cover). Timed: exactly ``steps`` steps, bracketed by device
    synchronisation. The run then continues to ``run_steps``; the whole run (first step to last,
    window included) is timed as well. The impactor depth of the reference's objPosition.dat
    (fluidHeight + r - y, IO.cpp:1323) is sampled every ``every`` steps."""
    lib = lbdem.Library.load()
    sim = lbdem.Simulation(lib, lbdem.params_from_values(vals, prec), ps, obs)
    t0 = time.time()
    sim.init()
    sim.synchronize()
    name = PREC_NAME[prec]
    log(f"[bench] {name}: init {time.time() - t0:.2f} s")
    fh, r_imp = float(vals["fluidHeight"]), float(obs.r[0])
    state = {"done": 0}
    depth_steps, depth, active_samples = [], [], []

    def advance(n):
        # in chunks that end on multiples of ``every``, sampling depth and active nodes there
        while n > 0:
            k = min(n, every - state["done"] % every)
            sim.cycle(k)
            state["done"] += k
            n -= k
            if state["done"] % every == 0:
                o = sim.objects()
                depth_steps.append(state["done"])
                depth.append(fh + r_imp - float(o[0, lbdem.O["X0"] + 1]))
                active_samples.append(sim.info()["nActive"])

    t_run0 = time.perf_counter()
    advance(warmup)
    if state["done"] < timed_from:
        advance(timed_from - state["done"])
    extra = 0
    while True:
        tu = sim.tuning()
        if (tu["tuned"] and tu["pair_captured"]) or extra >= 64:
            break
        advance(4)   # a call runs its first step alone, then captures or replays step pairs
        extra += 4
    if not (tu["tuned"] and tu["pair_captured"]):
        log(f"[bench] {name}: start-up not finished after {extra} extra steps: {tu}")
    sim.synchronize()
    start_step = state["done"]
    a0 = sim.info()["nActive"]
    mass0 = sim.total_mass()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.cycle(steps)   # steady-state step pairs replay a captured HIP graph
    sim.synchronize()
    t1 = time.perf_counter()
    state["done"] += steps
    inf = sim.info()
    a1 = inf["nActive"]
    mass = sim.total_mass()
    advance(max(0, run_steps - state["done"]))
    sim.synchronize()
    t_run1 = time.perf_counter()
    run_done = state["done"]
    tuning = sim.tuning()
    sim.close()
    fused_ms, launches, fused_active = fused_kernel_run(vals, ps, obs, prec, run_done, tuning["tile"][2], every)
    wall = t1 - t0
    active = 0.5 * (a0 + a1)
    run_wall = t_run1 - t_run0
    res = dict(prec=name, steps=steps, seconds=wall, ms_per_step=1e3 * wall / steps,
               active_cells=active, mlups_active=active * steps / wall / 1e6,
               mlups_box=inf["totNodes"] * steps / wall / 1e6,
               particle_steps_per_s=len(ps.r) * steps / wall,
               fused_ms_avg=fused_ms / launches, fused_launches=launches,
               total_mass=mass, total_mass_start=mass0, builds=inf["neighborListBuilds"],
               window_start_step=start_step, startup_extra_steps=extra, tile_z=tuning["tile"][2],
               run_steps=run_done, run_seconds=run_wall, run_ms_per_step=1e3 * run_wall / run_done,
               run_mlups_active=float(np.mean(active_samples)) * run_done / run_wall / 1e6 if active_samples else None,
               depth_steps=depth_steps, depth=depth)
    b = BYTES_PER_CELL[prec]
    res["fused_active_cells"] = fused_active
    res["fused_achieved_gbs"] = fused_active * b / (res["fused_ms_avg"] * 1e-3) / 1e9
    log(f"[bench] {name}: window steps {start_step}-{start_step + steps}: {res['ms_per_step']:.3f} ms/step, "
        f"{res['mlups_active']:.0f} MLUPS active; fused collide-stream {res['fused_ms_avg']:.3f} ms = "
        f"{res['fused_achieved_gbs']:.0f} GB/s; whole run of {run_done} steps {res['run_ms_per_step']:.3f} ms/step; "
        f"final depth {depth[-1] if depth else float('nan'):.4f}")
    return res


def fused_kernel_run(vals, ps, obs, prec, run_steps, tile_z, every):
    """The fused collide-stream kernel over a whole run: the same run again, every fused launch
    bracketed by HIP events on the library stream (launches outside graphs), with the z-walk the
    first run's tuner chose. Returns the summed kernel time, the launch count and the mean number of
    active nodes (sampled every ``every`` steps)."""
    lib = lbdem.Library.load()
    sim = lbdem.Simulation(lib, lbdem.params_from_values(vals, prec), ps, obs)
    sim.init()
    sim.set_tile(tile_z)
    sim.profile(True)
    act = []
    done = 0
    while done < run_steps:
        k = min(every, run_steps - done)
        sim.cycle(k)
        done += k
        act.append(sim.info()["nActive"])
    sim.synchronize()
    ms, n = sim.profile_read()
    sim.profile(False)
    sim.close()
    return ms, max(n, 1), float(np.mean(act))


def depth_error(res, ref):
    """Impactor depth of ``res`` against ``ref`` (the FP64 run): at the last common sample and at
    the deepest point of the FP64 curve, as relative errors (depth changes sign when the impactor
    meets the surface, so the error is taken where the FP64 depth is furthest from zero)."""
    n = min(len(res["depth"]), len(ref["depth"]))
    if n == 0:
        return None
    d, r = np.asarray(res["depth"][:n]), np.asarray(ref["depth"][:n])
    k = int(np.argmax(np.abs(r)))
    return {"final_step": int(ref["depth_steps"][n - 1]), "final_depth": round(float(d[-1]), 6),
            "final_depth_fp64": round(float(r[-1]), 6),
            "final_rel_err": float(abs(d[-1] - r[-1]) / max(abs(r[-1]), 1e-30)),
            "max_depth_step": int(ref["depth_steps"][k]), "max_depth_rel_err": float(abs(d[k] - r[k]) / max(abs(r[k]), 1e-30)),
            "max_abs_diff_over_run": float(np.abs(d - r).max())}


def cpu_baseline(steps=203, warm=3):
    """The reference program (oracle/_ref/ref_driver, compiled from the reference's sources) on a
    bounded sample of the reference-scale run (workloads.CONFIG0: 64^3 box, 480 spheres of radius
    3, impactor r 8): its first ``steps`` steps (of 2000), the first ``warm`` untimed, with as many
    OpenMP threads as the host gives one GPU (at most 16)."""
    drv = os.path.join(HERE, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        vals, ps, obs = workloads.impact_box(**workloads.CONFIG0)
        d = os.path.join(tmp, "case")
        os.makedirs(os.path.join(d, "input"))
        lbdem.write_getpot(os.path.join(d, "config.cfg"), vals)
        lbdem.write_particles(os.path.join(d, "input", "particles.dat"), ps.r, ps.x0 + 0.05)
        lbdem.write_objects(os.path.join(d, "input", "objects.dat"), obs.r, obs.x0, obs.x1)
        with open(os.path.join(d, "input", "body.dat"), "w") as fh:
            fh.write("0\n")
        nproc = os.cpu_count() or 1
        threads = max(1, min(nproc, 16))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), LBDEM_WARMUP=str(warm))
        t0 = time.time()
        r = subprocess.run([drv, "config.cfg", "out", str(steps), "dump"], cwd=d, env=env, capture_output=True,
                           text=True, timeout=600)
        if r.returncode:
            log(f"[bench] cpu baseline failed: {r.stdout[-500:]} {r.stderr[-500:]}")
            return None
        line = [l for l in r.stdout.splitlines() if l.startswith("LBDEM_TIMING")][-1]
        kv = dict(x.split("=") for x in line.split()[1:])
        sec = float(kv["seconds"])
        nst = int(kv["steps"])
        active = int(kv["active"])
        log(f"[bench] cpu baseline: {nst} steps in {sec:.2f} s ({time.time() - t0:.1f} s total), {active} active cells")
        return {"value": active * nst / sec / 1e6, "unit": "MLUPS", "cores": int(kv["threads"]), "kind": "reference",
                "host_cpus": nproc, "particle_steps_per_s": len(ps.r) * nst / sec,
                "sample": f"reference program (its sources compiled with g++ -O2 -fopenmp) on the reference-scale "
                          f"run: 64^3 box x/z periodic, fluid to 48, {len(ps.r)} spheres r=3 (phi~0.3), impactor r=8 "
                          f"at 56, vy=-0.005; steps {warm + 1}-{warm + nst} of 2000 timed ({warm} warm-up), "
                          f"{active} active cells, {kv['threads']} OpenMP threads of {nproc} host CPUs"}


def run_pool(lib, runs, prec, budget, max_concurrent, chunk, arrivals=None):
    """Submits ``runs`` to a run pool (all at once, or at ``arrivals`` seconds on the pool clock)
    and waits for all; returns per-run results and the pool statistics."""
    pool = lbdem.Pool(lib, budget_bytes=budget, max_concurrent=max_concurrent, steps_per_chunk=chunk)
    params = [lbdem.params_from_values(r.values, prec) for r in runs]
    ids = []
    t_base = pool.clock()
    for i, r in enumerate(runs):
        if arrivals is not None:
            wait = t_base + float(arrivals[i]) - pool.clock()
            if wait > 0:
                time.sleep(wait)
        ids.append(pool.submit(params[i], r.particles, r.objects, r.nsteps))
        if (i + 1) % 16 == 0:
            log(f"[bench] pool: {i + 1} runs submitted, stats {pool.stats()}")
    while True:  # a progress line every 20 s while the runs finish
        try:
            pool.drain(20.0)
            break
        except lbdem.LbdemError:
            log(f"[bench] pool: draining, stats {pool.stats()}")
    res = [pool.poll(rid) for rid in ids]
    st = pool.stats()
    pool.close()
    # A run may end with the reference's own TYPE ERROR (LB.cpp:1524-1527), e.g. when an impactor
    # sinks through the floor (objects have no wall contact in the reference DEM) and covers and
    # uncovers wall nodes; it ran all its steps and is timed like the others, and counted apart.
    bad = [x["message"] for x in res if x["state"] != lbdem.RUN_DONE and x["error"] != LBDEM_E_TYPE]
    if bad:
        raise RuntimeError(f"pool runs failed: {bad[:3]}")
    st["type_error_runs"] = sum(1 for x in res if x["state"] != lbdem.RUN_DONE)
    return res, st


def rates(res_like, span, prec):
    """Whole-leg rates of a set of runs finished in ``span`` seconds: box node updates, active node
    updates (each run's active nodes after its last step times its steps), particle steps, and the
    active rate as a fraction of the HBM roofline at the algorithmic bytes per active node."""
    node_steps = sum(x["nodeSteps"] for x, _ in res_like)
    active_steps = sum(float(x["nActive"]) * r.nsteps for x, r in res_like)
    mla = active_steps / span / 1e6
    return {"mlups_box": round(node_steps / span / 1e6, 1), "mlups_active": round(mla, 1),
            "particle_steps_per_s": round(sum(len(r.particles.r) * r.nsteps for _, r in res_like) / span, 1),
            "roofline_frac": round(mla * 1e6 * BYTES_PER_CELL[prec] / (PEAK_HBM_GBS * 1e9), 4)}


def batched(prec, n_boxes=64, nsteps=2000):
    """Many independent boxes of uneven size submitted together to the run pool: aggregate
    lattice updates per second over the whole batch."""
    lib = lbdem.Library.load()
    t0 = time.time()
    runs = workloads.batched_runs(n_boxes, nsteps)
    log(f"[bench] batched: {n_boxes} boxes generated in {time.time() - t0:.1f} s")
    est = [lbdem.estimate_bytes(lib, lbdem.params_from_values(r.values, prec), r.particles, r.objects) for r in runs]
    budget = int(1.05 * sum(est)) + (64 << 20)
    res, st = run_pool(lib, runs, prec, budget, max_concurrent=n_boxes, chunk=32)
    t_first = min(x["tSubmit"] for x in res)
    span = max(x["tDone"] for x in res) - t_first
    out = {"boxes": n_boxes, "steps_each": nsteps, "edges": sorted({r.edge for r in runs}),
           "total_nodes": int(sum(r.nodes for r in runs)), "seconds": round(span, 3)}
    out.update(rates(list(zip(res, runs)), span, prec))
    out.update({"hbm_budget_bytes": budget, "bytes_peak": st["bytesPeak"], "active_peak": st["activePeak"],
                "type_error_runs": st["type_error_runs"]})
    log(f"[bench] batched: {out}")
    return out


def served(prec, n_req, capacity_mlups, load=0.8, seed=20261204):
    """Poisson arrivals of impact runs at ``load`` x the capacity measured on the batched
    workload, served by the run pool; then the same requests, at the same arrival times, served one
    at a time with plain per-run launches (no pool, no graph replay) by a server that takes each
    request in arrival order once it has arrived and the previous one is done. Latency is arrival
    to result, measured on the host clock in both legs."""
    lib = lbdem.Library.load()
    mean_nodes = float(np.mean([e ** 3 for e in workloads.SERVE_EDGES]))
    mean_steps = 0.5 * (1000 + 8000)
    rate = load * capacity_mlups * 1e6 / (mean_nodes * mean_steps)
    t0 = time.time()
    arrivals, runs = workloads.served_requests(n_req, rate, seed=seed)
    log(f"[bench] served: {n_req} requests at {rate:.3f}/s generated in {time.time() - t0:.1f} s")
    budget = 16 << 30
    res, st = run_pool(lib, runs, prec, budget, max_concurrent=32, chunk=32, arrivals=arrivals)
    lat = np.array([x["tDone"] - x["tSubmit"] for x in res])
    t_first = min(x["tSubmit"] for x in res)
    span = max(x["tDone"] for x in res) - t_first
    node_steps = sum(x["nodeSteps"] for x in res)
    pool_out = {"p50_s": round(float(np.percentile(lat, 50)), 3), "p99_s": round(float(np.percentile(lat, 99)), 3),
                "max_s": round(float(lat.max()), 3), "mean_s": round(float(lat.mean()), 3), "makespan_s": round(span, 3),
                "runs_per_hour": round(3600.0 * n_req / span, 1)}
    pool_out.update(rates(list(zip(res, runs)), span, prec))
    pool_out.update({"active_peak": st["activePeak"], "bytes_peak": st["bytesPeak"],
                     "type_error_runs": st["type_error_runs"]})
    log(f"[bench] served (pool): {pool_out}")
    # plain sequential service, in real time
    seq_type_errors = 0
    lat_seq, service, seq_res = [], [], []
    t_base = time.perf_counter()
    for i, r in enumerate(runs):
        wait = float(arrivals[i]) - (time.perf_counter() - t_base)
        if wait > 0:
            time.sleep(wait)
        sim = lbdem.Simulation(lib, lbdem.params_from_values(r.values, prec), r.particles, r.objects)
        sim.use_graphs(False)
        t = time.perf_counter()
        sim.init()
        sim.cycle(r.nsteps)
        try:
            sim.synchronize()
            sim.objects()
            n_act = sim.info()["nActive"]
        except lbdem.LbdemError as e:  # the reference's TYPE ERROR (see run_pool): timed all the same
            if "TYPE ERROR" not in str(e):
                raise
            seq_type_errors += 1
            n_act = res[i]["nActive"]
        done = time.perf_counter()
        service.append(done - t)
        lat_seq.append(done - t_base - float(arrivals[i]))
        seq_res.append({"nodeSteps": res[i]["nodeSteps"], "nActive": n_act})
        sim.close()
        if (i + 1) % 8 == 0:
            log(f"[bench] served (plain sequential): {i + 1}/{n_req} runs")
    lat_seq = np.array(lat_seq)
    span_seq = time.perf_counter() - t_base - float(arrivals[0])
    seq_out = {"p50_s": round(float(np.percentile(lat_seq, 50)), 3), "p99_s": round(float(np.percentile(lat_seq, 99)), 3),
               "max_s": round(float(lat_seq.max()), 3), "mean_s": round(float(lat_seq.mean()), 3),
               "makespan_s": round(span_seq, 3), "runs_per_hour": round(3600.0 * n_req / span_seq, 1),
               "service_mlups_box": round(node_steps / sum(service) / 1e6, 1), "type_error_runs": seq_type_errors}
    seq_out.update(rates(list(zip(seq_res, runs)), span_seq, prec))
    log(f"[bench] served (plain sequential): {seq_out}")
    return {"requests": n_req, "rate_per_s": round(rate, 4), "offered_load": load,
            "capacity_mlups_box": round(capacity_mlups, 1), "edges": list(workloads.SERVE_EDGES),
            "steps_range": [1000, 8000], "hbm_budget_bytes": budget, "pool": pool_out, "plain_sequential": seq_out,
            "runs_per_hour_ratio": round(pool_out["runs_per_hour"] / seq_out["runs_per_hour"], 3),
            "p99_ratio_plain_over_pool": round(seq_out["p99_s"] / pool_out["p99_s"], 3)}


def precision_sweep(vals, ps, obs, checkpoints=(100, 1000, 5000), every=100):
    """The headline state run in FP64 (the comparison base), FP32-offset and FP16-storage
    populations side by side: impactor depth (its drop since step 0) every ``every`` steps, total
    fluid mass drift, and at each checkpoint the errors of the FP32/FP16 runs against FP64 over
    the nodes active in both (mass and density fields, maximum absolute difference over the FP64
    field's largest value) plus the largest particle position difference. Mass drift is measured
    from the end of the first step."""
    lib = lbdem.Library.load()
    precs = (lbdem.PREC_FP64, lbdem.PREC_FP32, lbdem.PREC_FP16)
    sims = {}
    for prec in precs:
        sm = lbdem.Simulation(lib, lbdem.params_from_values(vals, prec), ps, obs)
        sm.init()
        sims[prec] = sm
    y0 = {prec: float(sm.objects()[0, lbdem.O["X0"] + 1]) for prec, sm in sims.items()}
    # mass baseline after the first step (its coupling step takes the nodes inside spheres out of
    # the fluid mass, LB.cpp:388-395 counts them at initialisation)
    for sm in sims.values():
        sm.cycle(1)
    m0 = {prec: sm.total_mass() for prec, sm in sims.items()}
    depth = {PREC_NAME[p]: [] for p in precs}
    out = {"steps": [], "checkpoints": {}, "mass_baseline_step": 1}
    done = 1
    for target in range(every, max(checkpoints) + 1, every):
        for sm in sims.values():
            sm.cycle(target - done)
        for sm in sims.values():
            sm.synchronize()
        done = target
        out["steps"].append(done)
        for prec, sm in sims.items():
            depth[PREC_NAME[prec]].append(round(y0[prec] - float(sm.objects()[0, lbdem.O["X0"] + 1]), 6))
        if done in checkpoints:
            ref = sims[lbdem.PREC_FP64]
            t_ref, _, _ = ref.types()
            _, m_ref, n_ref = ref.fields(want_f=False)
            p_ref = ref.particles()[:, 0:3]
            d_ref = depth["fp64"][-1]
            cp = {}
            for prec in (lbdem.PREC_FP32, lbdem.PREC_FP16):
                sm = sims[prec]
                t, _, _ = sm.types()
                _, m, n = sm.fields(want_f=False)
                act = ((t_ref == 0) | (t_ref == 3)) & ((t == 0) | (t == 3))
                d = depth[PREC_NAME[prec]][-1]
                cp[PREC_NAME[prec]] = {
                    "impactor_depth": d, "impactor_depth_fp64": d_ref,
                    "impactor_depth_rel_err": abs(d - d_ref) / max(abs(d_ref), 1e-30),
                    "mass_field_err": float(np.abs(m[act] - m_ref[act]).max() / np.abs(m_ref[act]).max()),
                    "density_field_err": float(np.abs(n[act] - n_ref[act]).max() / np.abs(n_ref[act]).max()),
                    "node_type_mismatch": int(np.count_nonzero(t != t_ref)),
                    "particle_position_max_abs_diff": float(np.abs(sm.particles()[:, 0:3] - p_ref).max()),
                    "total_mass_drift_rel": (sm.total_mass() - m0[prec]) / m0[prec],
                }
            cp["fp64"] = {"impactor_depth": d_ref,
                          "total_mass_drift_rel": (ref.total_mass() - m0[lbdem.PREC_FP64]) / m0[lbdem.PREC_FP64]}
            out["checkpoints"][str(done)] = cp
            log(f"[bench] precision sweep step {done}: {cp}")
    out["impactor_depth"] = depth
    for sm in sims.values():
        sm.close()
    return out


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size (torch's copy kernel): reads and writes
    ``nbytes`` per launch, the reference for scaling the HBM byte counters of rocprofv3."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps, "ms_per_copy": round(dt * 1e3, 4),
                      "gbs_read_plus_write": round(2 * nbytes / dt / 1e9, 1)}), flush=True)


def batch_profile(prec, n_boxes=64, steps=400, tile=8):
    """The batched workload's boxes (64, edges 48-128) as one batch of the batch API (not the pool):
    ``steps`` steps replayed from graphs and timed on the host, then the same number launched one by
    one with a HIP event after each launch (lbdem_batch_profile_enable): the time of every batch
    kernel, summed over its launches, beside the replayed step."""
    lib = lbdem.Library.load()
    runs = workloads.batched_runs(n_boxes, 2000)
    sims = []
    for run in runs:
        s = lbdem.Simulation(lib, lbdem.params_from_values(run.values, prec), run.particles, run.objects)
        s.init()
        s.set_tile(tile)
        s.cycle(4)  # start-up steps alone
        sims.append(s)
    for s in sims:
        s.synchronize()
    nodes = sum(s.info()["totNodes"] for s in sims)
    batch = lbdem.Batch(lib, sims)
    batch.set_tile(tile)
    batch.cycle(20)  # records and captures the pair of each phase
    batch.synchronize()
    t0 = time.perf_counter()
    batch.cycle(steps)
    batch.synchronize()
    replay = (time.perf_counter() - t0) / steps
    active = sum(s.info()["nActive"] for s in sims)
    log(f"[bench] batch profile: graph replay {replay * 1e3:.3f} ms per step; timing launches")
    batch.profile(True)
    t0 = time.perf_counter()
    batch.cycle(steps)
    batch.synchronize()
    timed = (time.perf_counter() - t0) / steps
    rows = batch.profile_read()
    batch.profile(False)
    batch.close()
    for s in sims:
        s.close()
    tot = sum(r["ms"] for r in rows)
    rows.sort(key=lambda r: -r["ms"])
    out = {"batch_profile": {
        "boxes": n_boxes, "box_nodes": int(nodes), "active_nodes_end": int(active), "steps": steps, "tile_z": tile,
        "precision": PREC_NAME[prec],
        "ms_per_step_graph_replay": round(replay * 1e3, 4),
        "mlups_box_graph_replay": round(nodes / replay / 1e6, 1),
        "ms_per_step_launch_by_launch": round(timed * 1e3, 4),
        "kernel_ms_per_step_sum": round(tot / steps, 4),
        "kernels": [{"kernel": r["kernel"], "ms_per_step": round(r["ms"] / steps, 4),
                     "launches_per_step": round(r["launches"] / steps, 2),
                     "us_per_launch": round(1e3 * r["ms"] / max(r["launches"], 1), 2)} for r in rows]}}
    print(json.dumps(out), flush=True)


def step_chains(rows, first, last):
    """The dispatches between fused launches `first` and `last` of a kernel trace (rows: start ns,
    end ns, kernel name), one entry per step: the fused kernel's duration, the time to the next
    fused launch, and for each chain of the step (free surface with the hydrodynamic finalisation;
    DEM; coupling and edge marking) the span from its first start to its last end and its launch
    count."""
    rows = sorted(rows)
    fused = [r for r in rows if "k_fused" in r[2]]
    chain_of = (("free_surface", ("k_fs_", "k_hydro_finalize")),
                ("dem", ("k_dem_", "k_maxdisp", "k_cells_", "k_nbr_")))
    steps = []
    for i in range(first, min(last, len(fused) - 1)):
        f0, f1 = fused[i], fused[i + 1]
        spans = {}
        for r in rows:
            if f0[1] <= r[0] < f1[0]:
                ch = next((c for c, pre in chain_of if any(p in r[2] for p in pre)), "coupling_marking")
                s = spans.setdefault(ch, [r[0], r[1], 0])
                s[0], s[1], s[2] = min(s[0], r[0]), max(s[1], r[1]), s[2] + 1
        steps.append({"fused_us": (f0[1] - f0[0]) / 1e3, "to_next_fused_us": (f1[0] - f0[1]) / 1e3,
                      **{c: [(v[1] - v[0]) / 1e3, v[2]] for c, v in spans.items()}})
    return steps


def summarize_profile(paths, window=(1600, 1606)):
    """Reduce rocprofv3 CSV output (kernel trace or counter collection) to per-kernel means, written
    beside each CSV as <csv>.sum.json ("summary": kernel -> counter -> [launches, mean]; "per_launch":
    every value of the fused and copy kernels; for a kernel trace also "steps", the chains of the
    steps between fused launches window[0] and window[1] of the process); the CSV itself is removed,
    since a long run's per-dispatch rows run to hundreds of MB."""
    import collections
    import csv
    for path in paths:
        rows = []
        agg = collections.defaultdict(lambda: collections.defaultdict(lambda: [0, 0.0]))
        keep = collections.defaultdict(list)
        trace = "kernel_trace" in os.path.basename(path)
        with open(path) as fh:
            for r in csv.DictReader(fh):
                kn = r["Kernel_Name"]
                if trace:
                    rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), kn))
                    v, cn = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3, "dur_us"
                else:
                    v, cn = float(r["Counter_Value"]), r["Counter_Name"]
                a = agg[kn][cn]
                a[0] += 1
                a[1] += v
                if "k_fused" in kn or "copy" in kn.lower():
                    keep[kn + "|" + cn].append(v)
        names = list(agg)
        try:
            dm = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True).stdout.split("\n")
        except OSError:
            dm = []
        out = {"summary": {(dm[i] if i < len(dm) and dm[i] else n)[:200]: {c: [v[0], v[1] / v[0]] for c, v in agg[n].items()}
                           for i, n in enumerate(names)},
               "per_launch": {k[:40] + ".." + k[-60:]: v for k, v in keep.items()}}
        if trace:
            out["steps"] = step_chains(rows, *window)
        with open(path + ".sum.json", "w") as fh:
            json.dump(out, fh)
        os.remove(path)
        print("summarized", path, len(names), "kernels", flush=True)
        if trace:
            print(trace_tables(out))


def kernel_label(name):
    """k_name<args> from a (possibly mangled, truncated) kernel name: float / double / integer /
    bool template arguments in order."""
    import re
    m = re.search(r"(k_[a-z0-9_]+)", name)
    if not m:
        return name[:60]
    rest, args, i = name[m.end():], [], 1
    if rest.startswith("<"):
        return m.group(1) + rest[:rest.find(">") + 1]
    if not rest.startswith("I"):
        return m.group(1)
    while i < len(rest) and rest[i] != "E":
        if rest[i] in "fd":
            args.append("float" if rest[i] == "f" else "double")
            i += 1
        elif rest.startswith(("Li", "Lb"), i) and "E" in rest[i:]:
            j = rest.index("E", i)
            v = rest[i + 2:j]
            args.append(v if rest[i + 1] == "i" else ("true" if v == "1" else "false"))
            i = j + 1
        else:
            break
    return f"{m.group(1)}<{', '.join(args)}>"


def trace_tables(out, top=36):
    """Markdown tables of a summarised kernel trace: kernels by total time, and the steps' chains."""
    rows = sorted(((v["dur_us"][0] * v["dur_us"][1], k, v["dur_us"][0], v["dur_us"][1])
                   for k, v in out["summary"].items() if "dur_us" in v), reverse=True)
    tot = sum(r[0] for r in rows) or 1.0
    lines = ["| kernel | calls | total ms | avg us | % of GPU time |", "|---|---:|---:|---:|---:|"]
    for t, k, c, a in rows[:top]:
        lines.append(f"| `{kernel_label(k)}` | {c} | {t / 1e3:.2f} | {a:.1f} | {100 * t / tot:.1f} |")
    lines += ["", "| step (after fused launch) | fused kernel us | until the next fused launch us | free-surface chain us "
              "(launches) | DEM chain us (launches) | coupling + marking us (launches) |", "|---|---:|---:|---:|---:|---:|"]
    for i, st in enumerate(out.get("steps", [])):
        f, d, c = (st.get(k, [0.0, 0]) for k in ("free_surface", "dem", "coupling_marking"))
        lines.append(f"| {i + 1} | {st['fused_us']:.1f} | {st['to_next_fused_us']:.1f} | {f[0]:.1f} ({f[1]}) | "
                     f"{d[0]:.1f} ({d[1]}) | {c[0]:.1f} ({c[1]}) |")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16"])
    ap.add_argument("--run-steps", type=int, default=5000, help="length of each headline run (the config's 5000)")
    ap.add_argument("--timed-from", type=int, default=1500,
                    help="first step of the timed window (untimed steps before it; 1500 is in the impact phase)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--small", action="store_true", help="quarter-size box (for quick checks)")
    ap.add_argument("--batched", type=int, default=64, help="boxes in the batched run (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=2000)
    ap.add_argument("--served", type=int, default=48,
                    help="requests in the served run (0: skip); 48 puts the 99th percentile between the two "
                         "slowest results rather than on the slowest")
    ap.add_argument("--headline-only", action="store_true", help="skip the batched and served runs")
    ap.add_argument("--precision-sweep", type=int, default=0, metavar="STEPS",
                    help="only the precision sweep on the headline state (checkpoints 100, 1000, 5000 up to STEPS)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy three times (calibrates HBM byte counters)")
    ap.add_argument("--batch-profile", type=int, default=0, metavar="STEPS",
                    help="only time every kernel of the batched workload's boxes in one batch (HIP events)")
    ap.add_argument("--summarize-profile", nargs="+", metavar="CSV",
                    help="only reduce rocprofv3 CSV files to per-kernel means (<csv>.sum.json; removes the CSV)")
    ap.add_argument("--profile-window", type=int, nargs=2, default=[1600, 1606], metavar=("FIRST", "LAST"),
                    help="fused launches of a kernel trace whose steps --summarize-profile breaks down by chain")
    args = ap.parse_args()
    if args.summarize_profile:
        summarize_profile(args.summarize_profile, tuple(args.profile_window))
        return
    if args.batch_profile:
        batch_profile({"fp32": lbdem.PREC_FP32, "fp64": lbdem.PREC_FP64, "fp16": lbdem.PREC_FP16}[args.precision],
                      steps=args.batch_profile)
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the coupled step has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    prec = {"fp32": lbdem.PREC_FP32, "fp64": lbdem.PREC_FP64, "fp16": lbdem.PREC_FP16}[args.precision]
    wl = dict(workloads.HEADLINE)
    if args.small:
        wl.update(size=(128, 160, 128), fluid_height=128.0, n_spheres=3800, r_imp=8.0, y_imp=145.0)
    t0 = time.time()
    vals, ps, obs = workloads.impact_box(**wl)
    phi = workloads.volume_fraction(ps, wl["size"], wl["fluid_height"])
    log(f"[bench] packing {len(ps.r)} spheres phi={phi:.3f} in {time.time() - t0:.1f} s")
    if args.precision_sweep:
        cps = tuple(c for c in (100, 1000, 5000) if c <= args.precision_sweep) or (args.precision_sweep,)
        res = precision_sweep(vals, ps, obs, cps)
        print(json.dumps({"precision_sweep": res, "workload": "headline impact run (256x256x320, 31000 spheres)"}),
              flush=True)
        return
    rs = args.run_steps
    main_res = run_gpu(vals, ps, obs, prec, args.steps, args.warmup, run_steps=rs, timed_from=min(args.timed_from, rs))
    fp64 = fp16 = None
    if not args.no_fp64 and prec != lbdem.PREC_FP64:
        fp64 = run_gpu(vals, ps, obs, lbdem.PREC_FP64, args.steps, args.warmup, run_steps=rs,
                       timed_from=min(args.timed_from, rs))
    if not args.no_fp64 and prec == lbdem.PREC_FP32:  # the FP16-storage variant, measured beside it
        fp16 = run_gpu(vals, ps, obs, lbdem.PREC_FP16, args.steps, args.warmup, run_steps=rs,
                       timed_from=min(args.timed_from, rs))
    bat = srv = None
    if not args.headline_only and args.batched > 0:
        bat = batched(prec, args.batched, args.batched_steps)
    if not args.headline_only and args.served > 0:
        cap = bat["mlups_box"] if bat else main_res["mlups_box"]
        srv = served(prec, args.served, cap)
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    b = BYTES_PER_CELL[prec]
    ach = main_res["fused_achieved_gbs"]
    traffic = None
    if prec == lbdem.PREC_FP32 and not args.small and os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
        traffic = t["hbm_read_bytes_per_launch"] + t["hbm_write_bytes_per_launch"]
    out = {
        "metric": "coupled MLUPS at 256²×320 φ=0.5 vs FP64 plain; p50/p99 served-run latency",
        "value": round(main_res["mlups_active"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": PREC_NAME[prec],
        "data": "synthetic (seeded sphere packing and impactor, lattice units)",
        "config": {
            "workload": "impact run, box 256x256 horizontal x 320 vertical, fluid to 256, 31000 spheres r=4 "
                        f"(phi={phi:.3f}), impactor r=16 at 290, vy=-0.01, tau=0.8, x/z periodic, floor and lid walls"
                        if not args.small else "quarter-size impact run (128x128x160, 3800 spheres)",
            "precision": PREC_NAME[prec],
            "active_cells": int(main_res["active_cells"]),
            "mlups_box": round(main_res["mlups_box"], 1),
            "particle_steps_per_s": round(main_res["particle_steps_per_s"], 1),
            "fp64_plain_mlups": round(fp64["mlups_active"], 1) if fp64 else None,
            "speedup_vs_fp64_plain": round(main_res["mlups_active"] / fp64["mlups_active"], 3) if fp64 else None,
            "fp16_storage": {"mlups_active": round(fp16["mlups_active"], 1), "ms_per_step": round(fp16["ms_per_step"], 4),
                             "fused_kernel_ms": round(fp16["fused_ms_avg"], 4),
                             "fused_achieved_gbs": round(fp16["fused_achieved_gbs"], 1),
                             "bytes_per_active_node": BYTES_PER_CELL[lbdem.PREC_FP16],
                             "speedup_vs_fp64_plain": round(fp16["mlups_active"] / fp64["mlups_active"], 3) if fp64 else None}
                            if fp16 else None,
            "fused_kernel_ms": round(main_res["fused_ms_avg"], 4),
            "timed_window": {"first_step": main_res["window_start_step"], "steps": args.steps,
                             "untimed_steps_before": main_res["window_start_step"],
                             "startup_extra_steps": main_res["startup_extra_steps"],
                             "note": "untimed: --warmup steps, steps to the impact phase, then step pairs until the "
                                     "tile tuner is done and the step pair is captured"},
            "tile_z": main_res["tile_z"],
            "whole_run": {"steps": main_res["run_steps"], "seconds": round(main_res["run_seconds"], 3),
                          "ms_per_step": round(main_res["run_ms_per_step"], 4),
                          "mlups_active": round(main_res["run_mlups_active"], 1) if main_res["run_mlups_active"] else None,
                          "fp64_ms_per_step": round(fp64["run_ms_per_step"], 4) if fp64 else None,
                          "fp16_ms_per_step": round(fp16["run_ms_per_step"], 4) if fp16 else None},
            "impactor_depth_vs_fp64": depth_error(main_res, fp64) if fp64 else None,
            "fp16_impactor_depth_vs_fp64": depth_error(fp16, fp64) if (fp16 and fp64) else None,
            "neighbor_list_builds": main_res["builds"],
            "total_mass_drift_rel": (main_res["total_mass"] - main_res["total_mass_start"]) / main_res["total_mass_start"],
        },
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": int(main_res["fused_active_cells"] * b),
                     "kernel": "k_fused", "bytes_per_active_node": b,
                     "kernel_ms_avg": round(main_res["fused_ms_avg"], 4), "launches": main_res["fused_launches"],
                     "timing": "HIP events around every fused launch of a second, event-timed run of the whole "
                               "headline run (same z-walk); active nodes averaged over that run"},
        "cpu_baseline": cpu,
        "batched": bat,
        "served": srv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
