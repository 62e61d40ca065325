"""The FP64 CPU oracle against golden fixtures written by the reference program itself
(tests/golden/make_golden.py): node types, particle flags, solid indices, populations, masses,
particle / impactor / wall records, Verlet lists, total fluid mass and, where the fixture holds
them, the stress diagnostics must agree bit for bit."""
import json
import os

import numpy as np
import pytest

import cases

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "manifest.json")) as fh:
    MANIFEST = json.load(fh)

# group fixtures by case so each case runs once
BY_CASE = {}
for key, meta in MANIFEST.items():
    BY_CASE.setdefault(meta["case"], []).append(meta["step"])


@pytest.mark.parametrize("name", sorted(BY_CASE))
def test_oracle_matches_reference_fixtures(oracle_lib, case_root, name):
    d = cases.make_case(case_root, name)
    sim = cases.open_sim(oracle_lib, d, stress=True)
    done = 0
    for step in sorted(BY_CASE[name]):
        sim.cycle(step - done)
        done = step
        g = np.load(os.path.join(GOLD, f"{name}_s{step}.npz"), allow_pickle=False)
        st = cases.state(sim, stress=True)
        assert np.array_equal(st["type"], g["type"]), f"{name} step {step}: node types"
        assert np.array_equal(st["pflag"], g["pflag"]), f"{name} step {step}: particle flags"
        assert np.array_equal(st["solid"], g["solid"]), f"{name} step {step}: solid indices"
        act = g["active"]
        assert np.array_equal(st["f"][:, act], g["f_active"]), f"{name} step {step}: populations"
        assert np.array_equal(st["mass"][act], g["mass_active"]), f"{name} step {step}: mass"
        assert np.array_equal(st["n"][act], g["n_active"]), f"{name} step {step}: density"
        assert np.array_equal(st["particles"], g["particles"]), f"{name} step {step}: particles"
        assert np.array_equal(st["objects"], g["objects"]), f"{name} step {step}: objects"
        assert np.array_equal(st["walls"], g["walls"]), f"{name} step {step}: walls"
        if "bodies" in g:
            assert np.array_equal(st["bodies"], g["bodies"]), f"{name} step {step}: bodies"
        assert np.array_equal(st["ncount"], g["ncount"]), f"{name} step {step}: neighbour counts"
        cnt = g["ncount"]
        for p in range(len(cnt)):
            assert np.array_equal(st["nlist"][p, :cnt[p]], g["nlist"][p, :cnt[p]]), f"{name} step {step}: list of {p}"
        assert st["totalMass"] == float(g["totalMass"]), f"{name} step {step}: total fluid mass"
        assert st["info"]["totalMassInit"] == float(g["totalMassInit"])
        if "stress" in g:
            for k in ("stress", "stotal", "disp", "pstress"):
                if k in g:
                    assert np.array_equal(st[k], g[k]), f"{name} step {step}: {k}"


def test_fixtures_exercise_every_event(oracle_lib, case_root):
    """The suspension fixture window contains contacts, lubrication, impactor contact and
    lubrication, wall contacts and interface conversions (otherwise the bit-exact check above
    would leave those paths unpinned)."""
    import lbdem
    d = cases.make_case(case_root, "suspension")
    sim = cases.open_sim(oracle_lib, d)
    ev = dict(contact=0, lub=0, objc=0, objl=0, wallc=0, conv=0)
    for _ in range(120):
        sim.cycle(1)
        p, o, w, inf = sim.particles(), sim.objects(), sim.walls(), sim.info()
        ev["contact"] += int((np.abs(p[:, lbdem.P["FCONTACT"]:lbdem.P["FCONTACT"] + 3]).sum(1) > 0).sum())
        ev["lub"] += int((np.abs(p[:, lbdem.P["FLUB"]:lbdem.P["FLUB"] + 3]).sum(1) > 0).sum())
        ev["objc"] += int(np.abs(o[0, lbdem.O["FCONTACT"]:lbdem.O["FCONTACT"] + 3]).sum() > 0)
        ev["objl"] += int(np.abs(o[0, lbdem.O["FLUB"]:lbdem.O["FLUB"] + 3]).sum() > 0)
        ev["wallc"] += int((np.abs(w[:, lbdem.W["FCONTACT"]:lbdem.W["FCONTACT"] + 3]).sum(1) > 0).sum())
        ev["conv"] += int(inf["lastSurplus"] != 0.0)
    assert all(v > 0 for v in ev.values()), ev


def test_cell_fixtures_cross_the_seams():
    """The "cells" fixtures cover what the link-cell branch must get right: 6 cells per axis, a
    sphere that crossed the x seam and one that crossed the z seam between steps 1 and 26 (the
    DEM periodic wrap, DEM.cpp:2769-2836), and listed pairs whose raw separation spans the box (the
    wrapped walk, DEM.cpp:585-603); the thin-z fixtures hold repeated partners (2 and 1 z-cells)
    and empty lists (0 z-cells)."""
    import lbdem
    s1 = np.load(os.path.join(GOLD, "cells_s1.npz"))
    s26 = np.load(os.path.join(GOLD, "cells_s26.npz"))
    x1, x26 = s1["particles"][:, 0:3], s26["particles"][:, 0:3]
    assert x1[452, 0] > 63.0 and x26[452, 0] < 1.0          # across x = 64 -> 0
    assert x1[454, 2] < 1.0 and x26[454, 2] > 63.0          # across z = 0 -> 64
    assert np.all(s1["particles"][:, lbdem.P["R"]] == 4.0)   # cut 10 over a 64^3 DEM box: 6 cells
    cnt, lst = s26["ncount"], s26["nlist"]
    span = [abs(x26[a] - x26[b]).max() for a in range(len(cnt)) for b in lst[a, :cnt[a]]]
    assert max(span) > 32.0
    for name, dup in (("thin_z2_s40", True), ("thin_z1_s40", True)):
        g = np.load(os.path.join(GOLD, f"{name}.npz"))
        rows = [g["nlist"][a, :g["ncount"][a]] for a in range(len(g["ncount"]))]
        assert any(len(r) != len(set(r.tolist())) for r in rows) == dup, name
    assert np.all(np.load(os.path.join(GOLD, "thin_z0_s20.npz"))["ncount"] == 0)


def test_rheology_fixtures_cover_the_viscosity_range(oracle_lib, case_root):
    """The rheology fixtures exercise the apparent viscosity on both sides of its clamp: after 40
    steps the node viscosities vary with the shear rate, and the BINGHAM, FRICTIONAL and VOELLMY
    cases hold nodes both below and at maxVisc = (1.8 - 0.5) / 3 (lattice.h:27); the Smagorinsky
    case stays above its Newtonian viscosity 0.1 by the turbulent term."""
    top = (1.8 - 0.5) / 3.0
    for name in cases.RHEOLOGY_CASES:
        sim = cases.open_sim(oracle_lib, cases.make_case(case_root, name))
        sim.cycle(39)
        sim.store_velocity()
        sim.cycle(1)
        _, visc = sim.velocity()
        v = visc[visc > 0]
        assert v.size > 100 and v.max() - v.min() > 1e-4, name
        if name == "smagorinsky":
            assert v.min() >= 0.1 and v.max() < top, name
        else:
            assert (v >= top).any() and (v < 0.99 * top).sum() > 10, name
        sim.close()


def test_slip_fixture_pins_the_slip_term(oracle_lib, case_root):
    """The slip-wall fixture depends on the slip term: the same case with slipCoefficient 0 (every
    wall link bounces back) ends 40 steps later with populations well apart from the fixture's."""
    g = np.load(os.path.join(GOLD, "slip_walls_s40.npz"), allow_pickle=False)
    d = cases.make_case(case_root, "slip_walls", overrides={"slipCoefficient": 0.0}, subdir="slip_zero")
    sim = cases.open_sim(oracle_lib, d)
    sim.cycle(40)
    st = cases.state(sim)
    act = g["active"]
    assert np.abs(st["f"][:, act] - g["f_active"]).max() > 1e-6
    sim.close()
    # moving slip walls differ from static ones only in the free-surface mass exchange of their
    # slip links (LB.cpp:1595-1625): the two fixtures share the populations' reflection and part
    # in the interface masses
    h = np.load(os.path.join(GOLD, "slip_dyn_walls_s40.npz"), allow_pickle=False)
    assert float(h["totalMass"]) != float(g["totalMass"])
