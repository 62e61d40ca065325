"""Closed-form and invariant checks of the FP64 oracle that do not depend on the reference run:
D3Q19 equilibrium moments, exact mass conservation of BGK + Guo forcing with half-way
bounce-back in a closed box, a flat free surface at rest, and the symmetric Verlet list."""
import numpy as np
import pytest

import cases
import lbdem

W = np.array([12 / 36] + [2 / 36] * 6 + [1 / 36] * 12)
C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [-1, 1, 0], [1, -1, 0], [0, 1, 1], [0, -1, -1], [0, -1, 1], [0, 1, -1],
              [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1]], dtype=float)


def test_lattice_tables_closed_form():
    assert abs(W.sum() - 1.0) < 1e-15
    assert np.allclose(W @ C, 0.0, atol=1e-16)
    assert np.allclose((W[:, None, None] * C[:, :, None] * C[:, None, :]).sum(0), np.eye(3) / 3.0, atol=1e-16)


def _closed_box(root, **kw):
    v, *_ = cases._case_values("lb_closed")
    v.update(kw)
    d = cases.make_case(root, "lb_closed")
    lbdem.write_getpot(f"{d}/config.cfg", v)
    return d


def test_initial_equilibrium_moments(oracle_lib, case_root):
    d = _closed_box(case_root, initVelocityX=0.01, initVelocityZ=-0.02, lbFX=1e-4, lbFY=0.0)
    sim = cases.open_sim(oracle_lib, d)
    t, _, _ = sim.types()
    f, _, _ = sim.fields()
    act = t == 0
    fa = f[:, act]
    n = fa.sum(0)
    u = (C.T @ fa) / n
    # node::initialize: n = rho0, u = lbF/(2 n) + u0
    assert np.allclose(n, 1.0, atol=1e-15)
    assert np.allclose(u[0], 1e-4 / 2 + 0.01, atol=1e-15)
    assert np.allclose(u[2], -0.02, atol=1e-15)
    # second moment of the equilibrium: sum f c c = n (u u + I/3)
    P = np.einsum("qi,qj,qn->ijn", C, C, fa)
    expect = n * (u[:, None, :] * u[None, :, :]) + np.eye(3)[:, :, None] * n / 3
    assert np.allclose(P, expect, atol=1e-15)


def test_closed_box_conserves_mass(oracle_lib, case_root):
    d = _closed_box(case_root, lbFX=2e-4, lbFY=-1e-4)
    sim = cases.open_sim(oracle_lib, d)
    t, _, _ = sim.types()
    f0, _, _ = sim.fields()
    m0 = f0[:, t == 0].sum()
    sim.cycle(1)
    f1, _, _ = sim.fields()
    # the body force first accelerates the fluid along +x
    assert (C.T @ f1[:, t == 0]).sum(1)[0] > (C.T @ f0[:, t == 0]).sum(1)[0]
    sim.cycle(39)
    f1, _, _ = sim.fields()
    assert abs(f1[:, t == 0].sum() - m0) < 1e-11 * m0


def test_flat_surface_at_rest_is_steady(oracle_lib, case_root):
    v, *_ = cases._case_values("fs_impactor")
    v.update(lbFY=0.0, objWaitTime=0)
    d = cases.make_case(case_root + "/rest", "fs_impactor")
    lbdem.write_getpot(f"{d}/config.cfg", v)
    lbdem.write_objects(f"{d}/input/objects.dat", [], np.zeros((0, 3)), np.zeros((0, 3)))
    sim = cases.open_sim(oracle_lib, d)
    t0, _, _ = sim.types()
    m0 = sim.total_mass()
    for _ in range(30):
        sim.cycle(1)
        assert sim.info()["lastSurplus"] == 0.0
    t1, _, _ = sim.types()
    assert np.array_equal(t0, t1)
    assert abs(sim.total_mass() - m0) < 1e-12 * m0


def test_verlet_list_is_symmetric_and_complete(oracle_lib, case_root):
    d = cases.make_case(case_root, "suspension")
    sim = cases.open_sim(oracle_lib, d)
    cnt, lst = sim.neighbors()
    p = sim.particles()
    x, r = p[:, 0:3], p[:, lbdem.P["R"]]
    n = len(r)
    for a in range(n):
        partners = set(lst[a, :cnt[a]].tolist())
        for b in range(n):
            if a == b:
                continue
            dx = x[a] - x[b]
            inside = (dx @ dx) <= (r[a] + r[b] + 2.0) ** 2
            assert (b in partners) == inside
            assert (a in set(lst[b, :cnt[b]].tolist())) == (b in partners)



@pytest.mark.parametrize("name", cases.CELL_CASES)
def test_cell_branch_lists_equal_brute_force(oracle_lib, case_root, name):
    """The oracle's link-cell neighbour list (DEM.cpp:556-631) against a brute-force pair search
    with the multiplicity of the wrapped 27-cell walk stated in closed form
    (cases.brute_neighbours): equal partner multisets at initialisation and after every rebuild of
    a 30-step window (for "cells" the window has spheres on both sides of both periodic seams, and
    the four spheres above the surface are listed with bed spheres through the y image of the DEM
    box)."""
    vals, *_ = cases._case_values(name)
    d = cases.make_case(case_root, name)
    sim = cases.open_sim(oracle_lib, d)
    L = cases.dem_size(vals)
    R = sim.particles()[:, lbdem.P["R"]]
    # a rebuild happens inside a step's DEM pass, from the positions the previous step left
    x_before = sim.particles()[:, 0:3]
    builds = sim.info()["neighborListBuilds"]
    checked = 0
    for step in range(31):
        if step:
            sim.cycle(1)
            now = sim.info()["neighborListBuilds"]
            if now == builds:
                x_before = sim.particles()[:, 0:3]
                continue
            builds = now
        expect = cases.brute_neighbours(x_before, R, L)
        cnt, lst = sim.neighbors()
        for a in range(len(cnt)):
            assert sorted(lst[a, :cnt[a]].tolist()) == expect[a], f"{name} step {step}: partners of {a}"
        checked += 1
        x_before = sim.particles()[:, 0:3]
    assert checked >= 2, checked
