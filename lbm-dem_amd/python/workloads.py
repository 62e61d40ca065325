"""Synthetic impact-run workloads in the reference's input terms (reference axes: y vertical).

``impact_box`` builds the getpot values, particles and impactor of one run:
  * a box with no-slip floor and lid (y faces) and either no-slip walls on the x and z faces (the
    reference example's geometry) or periodic x and z faces (``periodic=True``); lattice units
    (unitLength = unitTime = unitDensity = 1), tau = 0.8 (initVisc = 0.1), gravity lbFY = -1e-5;
  * fluid up to ``fluid_height`` with a free surface;
  * ``n_spheres`` spheres of radius ``r`` from a seeded random-sequential-addition + growth /
    relaxation packing below the surface (lbm-dem_amd/host/packing.cpp);
  * one impactor of radius ``r_imp`` at (x/2, y_imp, z/2) moving with vy = -speed, density 3;
  * sphere density 2.5 (the reference's explicit interior forcing, LB.cpp:1901-1978 used one
    step late by the DEM, amplifies velocity differences when spheres are as dense as the fluid).
DEM parameters in lattice units (dt = 1): linearStiff 2.0 (contact time sqrt(m/k) ~ 12 steps for
r = 4), friction 0.3, lubrication with kinVisc = 0.1 (the fluid's), lubNormal 1.0,
lubTangential 0.5, lubRotation 0.4, lubCutoff 0.2, electrostatic repulsion 0.01 with Debye
length 0.5, objWaitTime 0 (the impactor moves from the first step).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Dict, Tuple

import numpy as np

import lbdem

TOOLS = os.path.join(lbdem.PKG, "lib", "liblbdem_tools.so")


def pack(n: int, r: float, lo, hi, seed: int, relax: int = 500) -> np.ndarray:
    lib = C.CDLL(TOOLS)
    x = np.zeros(3 * n)
    mo = C.c_double()
    rc = lib.lbdem_pack_spheres(C.c_int(n), C.c_double(r), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi),
                                C.c_ulonglong(seed), C.c_int(relax), x.ctypes.data_as(C.POINTER(C.c_double)),
                                C.byref(mo))
    if rc:
        raise RuntimeError(f"packing of {n} spheres of radius {r} failed")
    return x.reshape(n, 3)


def impact_box(size=(256, 320, 256), fluid_height=256.0, n_spheres=31000, r=4.0, r_imp=16.0, y_imp=290.0,
               speed=0.01, seed=20261202, relax=500, periodic=False
               ) -> Tuple[Dict[str, object], lbdem.ParticleSet, lbdem.ObjectSet]:
    X, Y, Z = size
    side = 4 if periodic else 7
    v = dict(lbdem.DEFAULTS)
    v.update({
        "demSolve": 1, "lbSolve": 1, "freeSurfaceSolve": 1, "forceFieldSolve": 1, "demInitialRepeat": 0.0,
        "maxTime": 5000.0, "unitLength": 1.0, "unitTime": 1.0, "unitDensity": 1.0,
        "lbSizeX": float(X), "lbSizeY": float(Y), "lbSizeZ": float(Z),
        "rheologyModel": "NEWTONIAN", "initVisc": 0.1, "lbFY": -1.0e-5, "fluidHeight": float(fluid_height),
        "boundary0": side, "boundary1": side, "boundary2": 7, "boundary3": 7, "boundary4": side, "boundary5": side,
        "density": 2.5, "objectDensity": 3.0, "bodyDensity": 12.0, "linearStiff": 2.0, "frictionCoefPart": 0.3,
        "lubNormal": 1.0, "lubTangential": 0.5, "lubRotation": 0.4, "lubCutoff": 0.2, "kinVisc": 0.1,
        "electroForce": 0.01, "debyeLength": 0.5, "objWaitTime": 0,
    })
    ps = lbdem.ParticleSet()
    if n_spheres > 0:
        lo = (1.0, 1.0, 1.0)
        hi = (X - 1.0, fluid_height, Z - 1.0)
        ps.x0 = pack(n_spheres, r, lo, hi, seed, relax)
        ps.r = np.full(n_spheres, r)
        ps.size = np.ones(n_spheres, np.int32)
    obs = lbdem.ObjectSet(r=np.array([r_imp]), x0=np.array([[X / 2.0, y_imp, Z / 2.0]]),
                          x1=np.array([[0.0, -speed, 0.0]]))
    return v, ps, obs


def volume_fraction(ps: lbdem.ParticleSet, size, fluid_height) -> float:
    X, Y, Z = size
    return float(np.sum(4.0 / 3.0 * math.pi * ps.r ** 3) / ((X - 2.0) * (fluid_height - 1.0) * (Z - 2.0)))


# The headline workload: 256 x 256 x 320 box (x, z horizontal and periodic; y vertical up to 320,
# floor and lid walls), fluid to 256, 31 000 spheres of radius 4 (phi ~ 0.5), impactor r = 16 at
# height 290, v = -0.01.
HEADLINE = dict(size=(256, 320, 256), fluid_height=256.0, n_spheres=31000, r=4.0, r_imp=16.0, y_imp=290.0,
                speed=0.01, seed=20261202, periodic=True)
