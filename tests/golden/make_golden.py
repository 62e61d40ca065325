"""Writes the golden fixtures in this directory from the reference program.

Runs the reference compiled from its own sources (oracle/_ref/ref_driver, built by
``make -C oracle ref`` in a checkout that has the reference at /root/reference) on the small
cases of tests/cases.py, single-threaded, and stores the solver state after selected steps as
compressed .npz files: node types / particle flags / solid indices, post-streaming populations
of active nodes, density and mass fields, particle / object / wall records, Verlet lists and
the total fluid mass. Also runs the reference's own main loop on the host-program cases
(cases.HOST_RUNS) and keeps the result files it writes (objPosition.dat, objForce.dat,
wallForce.dat) under host_<case>/, the fluidRecycle files of the recycle run
(cases.RECYCLE_RUN) under recycle_<case>/ (gzip), which the "restart" case reads back, and the
.vti / .vtu files of the fluid run (cases.FLUID_RUN) under fluid_<case>/. Only data is stored.

Cases larger than SAMPLE_ABOVE nodes keep the populations, density and mass of a seeded sample
of active nodes (``sample`` holds their indices) instead of every active node; everything else
is stored whole.

``--inputs`` writes the sphere centres of the packed cases (cases.PACKED) to inputs_<case>.npy
from the seeded packing tool (lbm-dem_amd/lib/liblbdem_tools.so); they are committed, so the
cases do not depend on the tool.

usage: python tests/golden/make_golden.py [--inputs] [--host-only] [--cases name,name,...]
(--cases: only the named cases of the plan, merged into the existing manifest)
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import cases  # noqa: E402

DRIVER = os.path.join(cases.REPO, "oracle", "_ref", "ref_driver")
PLAN = {"lb_closed": [1, 12], "fs_impactor": [1, 40], "suspension": [1, 40, 120], "periodic": [40], "units": [40],
        "moving": [1, 30], "startup": [5, 30], "two_impactors": [1, 40],
        "cells": [1, 26, 60], "thin_z2": [1, 40], "thin_z1": [40], "thin_z0": [20], "restart": [0, 1, 20],
        "bingham": [1, 40], "frictional": [1, 40], "voellmy": [1, 40], "smagorinsky": [1, 12],
        "body_wca": [1, 40], "body_cubic": [1, 40], "dyn_walls": [1, 40], "wall_codes": [1, 40],
        "slip_walls": [1, 40], "slip_dyn_walls": [1, 40]}
SAMPLE_ABOVE = 8_000   # nodes
SAMPLE_NODES = 1536


def make_inputs():
    """Sphere centres (particles.dat columns, i.e. before the reader's -0.05) of the packed cases."""
    sys.path.insert(0, os.path.join(cases.REPO, "lbm-dem_amd", "python"))
    import workloads
    for name in cases.PACKED:
        v, size, fh, r, n, seed, *_ = cases.packed_values(name)
        X, Y, Z = size
        if name == "cells":
            x0 = workloads.pack(n, r, (1.0, 1.0, 1.0), (X - 1.0, fh, Z - 1.0), seed, relax=300)
            # move the packing by 30 lattice units along x and z (periodic) so spheres sit on both
            # sides of the seams, then add the seam-crossing pairs
            x0[:, 0] = np.mod(x0[:, 0] + 30.0, float(X))
            x0[:, 2] = np.mod(x0[:, 2] + 30.0, float(Z))
            pc = np.concatenate([x0 + 0.05, cases.SEAM_PAIRS])
        else:
            x0 = workloads.pack(n, r, (1.0, 1.0, 0.5), (X - 1.0, fh - 1.0, Z - 0.5), seed, relax=200)
            pc = x0 + 0.05
        out = os.path.join(HERE, f"inputs_{name}.npy")
        np.save(out, pc)
        print("inputs of", name, pc.shape, "->", out)


def host_outputs():
    """Result files of the reference's main loop (IO::outputStep) on the host-program cases."""
    import shutil
    with tempfile.TemporaryDirectory() as tmp:
        for name, over in cases.HOST_RUNS.items():
            d = cases.make_case(tmp, name, over)
            steps = int(over["maxTime"]) + 1
            env = dict(os.environ, OMP_NUM_THREADS="1")
            r = subprocess.run([DRIVER, "config.cfg", "out", str(steps), "dump"], cwd=d, env=env, capture_output=True,
                               text=True)
            if r.returncode:
                sys.exit(f"reference run failed for {name}:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
            dst = os.path.join(HERE, f"host_{name}")
            os.makedirs(dst, exist_ok=True)
            for f in cases.HOST_FILES:
                shutil.copy(os.path.join(d, "out", f), os.path.join(dst, f))
            print("host outputs of", name, "->", dst)


def fluid_files(run, prefix, only=None, sub="fluidData", starts=None):
    """Runs the reference's main loop on a file run (case, overrides) for as many steps as its own
    loop would take, and keeps the files it writes into <sub>/ (all of them, those named in ``only``
    or those whose names start with one of ``starts``) gzip-compressed under <prefix>_<case>/."""
    import gzip
    name, over = run
    with tempfile.TemporaryDirectory() as tmp:
        d = cases.make_case(tmp, name, over)
        vals, *_ = cases._case_values(name)
        vals.update(over)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([DRIVER, "config.cfg", "out", str(cases.host_steps(vals)), "dump"], cwd=d, env=env,
                           capture_output=True, text=True)
        if r.returncode:
            sys.exit(f"reference run failed for {prefix}_{name}:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        src_dir = os.path.join(d, "out", sub)
        dst = os.path.join(HERE, f"{prefix}_{name}")
        os.makedirs(dst, exist_ok=True)
        names = sorted(os.listdir(src_dir)) if only is None else list(only)
        if starts is not None:
            names = [f for f in names if f.startswith(tuple(starts))]
        for f in names:
            with open(os.path.join(src_dir, f), "rb") as a, gzip.GzipFile(os.path.join(dst, f + ".gz"), "wb",
                                                                         mtime=0) as b:
                b.write(a.read())
        print(f"{prefix} files of", name, "->", dst, names)


def recycle_outputs():
    """The fluidRecycle files the reference writes on the recycle run (IO.cpp:203-210,
    973-997), gzip-compressed under recycle_<case>/ (the restart case reads one back), the DEM VTK
    files of the DEM run under dem_<case>/, and the Eulerian / Lagrangian fluid files of the fluid
    run under fluid_<case>/."""
    fluid_files(cases.RECYCLE_RUN, "recycle", [cases.recycle_name(s) for s in cases.RECYCLE_STEPS])
    fluid_files(cases.DEM_RUN, "dem", sub="particleData", starts=cases.DEM_FILES)
    fluid_files(cases.FLUID_RUN, "fluid")


def main():
    if not os.path.exists(DRIVER):
        sys.exit(f"{DRIVER} missing: run `make -C oracle ref` where the reference sources are present")
    if "--inputs" in sys.argv:
        make_inputs()
        return
    only = None
    if "--cases" in sys.argv:
        only = sys.argv[sys.argv.index("--cases") + 1].split(",")
        unknown = [n for n in only if n not in PLAN]
        if unknown:
            sys.exit(f"not in the plan: {unknown}")
    else:
        host_outputs()
        recycle_outputs()
    if only is not None and "restart" in only:
        recycle_outputs()
    if "--host-only" in sys.argv:
        return
    manifest = {}
    if only is not None:
        with open(os.path.join(HERE, "manifest.json")) as fh:
            manifest = json.load(fh)
    with tempfile.TemporaryDirectory() as tmp:
        for name, steps in PLAN.items():
            if only is not None and name not in only:
                continue
            d = cases.make_case(tmp, name)
            env = dict(os.environ, OMP_NUM_THREADS="1")
            r = subprocess.run([DRIVER, "config.cfg", "out", str(max(steps)), "dump"] + [str(s) for s in steps],
                               cwd=d, env=env, capture_output=True, text=True)
            if r.returncode:
                sys.exit(f"reference run failed for {name}:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
            for s in steps:
                dump = cases.read_dump(os.path.join(d, f"dump_{s}.bin"))
                act = (dump["type"] == 0) | (dump["type"] == 3)
                out = os.path.join(HERE, f"{name}_s{s}.npz")
                sel = np.nonzero(act)[0]
                if dump["N"] > SAMPLE_ABOVE:
                    rng = np.random.default_rng(1000 + s)
                    sel = np.sort(rng.choice(sel, size=min(SAMPLE_NODES, sel.size), replace=False))
                extra = {k: dump[k] for k in ("stress", "stotal", "disp", "pstress") if k in dump}
                np.savez_compressed(
                    out, **extra, type=dump["type"], pflag=dump["pflag"], solid=dump["solid"], active=sel,
                    sampled=np.array(dump["N"] > SAMPLE_ABOVE),
                    f_active=dump["f"][:, sel], n_active=dump["n"][sel], mass_active=dump["mass"][sel],
                    particles=dump["particles"], objects=dump["objects"], walls=dump["walls"],
                    ncount=dump["ncount"], nlist=dump["nlist"], totalMass=dump["totalMass"], bodies=dump["bodies"],
                    totalMassInit=dump["totalMassInit"], size=np.array(dump["size"]))
                manifest[f"{name}_s{s}"] = dict(case=name, step=s, size=list(dump["size"]),
                                                 active=int(act.sum()), bytes=os.path.getsize(out))
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main()
