"""ctypes binding of the lbdem C ABI (include/lbdem.h) plus the reference's input formats.

The same binding drives any library exporting the ABI (the MI355X library
``lbm-dem_amd/lib/liblbdem.so``; tests also point it at the CPU oracle).

Input formats follow the reference:
  * getpot config  ``name = value  # comment`` (getpot.h:110-117, macros.h:14-34),
    keys and defaults as read by LB::latticeBoltzmannGet (LB.cpp:128-338),
    DEM::discreteElementGet (DEM.cpp:17-75) and parseConfigFile (main.cpp:128-185);
  * ``input/particles.dat``: count, then rows ``index size r x y z vx vy vz wx wy wz q0..q3 q1(4)``
    read with ``x0 = x*scale - 0.05`` and ``r *= scale`` (DEM.cpp:81-134);
  * ``input/objects.dat``: count, then rows ``index r x y z vx vy vz`` (DEM.cpp:136-165).
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "liblbdem.so")

PREC_FP64, PREC_FP32, PREC_FP16 = 0, 1, 2
RHEO_NEWTONIAN, RHEO_BINGHAM, RHEO_FRICTIONAL, RHEO_VOELLMY = 0, 1, 2, 3  # include/lbdem.h LBDEM_RHEO_*
Q = 19

# record layouts (include/lbdem.h)
P = dict(X0=0, X1=3, W1=6, FHYDRO=9, MHYDRO=12, FCONTACT=15, MCONTACT=18, FLUB=21, MLUB=24, FREP=27,
         FTOTAL=30, MTOTAL=33, FGRAV=36, FLUIDVOL=39, XIW=40, XIO=43, R=46, M=47, REC=48)
O = dict(X0=0, X1=3, W1=6, FHYDRO=9, MHYDRO=12, FCONTACT=15, MCONTACT=18, FLUB=21, MLUB=24, FTOTAL=27,
         MTOTAL=30, FGRAV=33, R=36, M=37, REC=38)
W = dict(FHYDRO=0, FCONTACT=3, FLUB=6, FTOTAL=9, REC=12)
B = dict(X0=0, X1=3, FSPRING=6, FTOTAL=9, FGRAV=12, L=15, L0=16, K=17, R=18, M=19, REC=20)  # LBDEM_B_*
MAX_BODIES = 8


class Params(C.Structure):
    _fields_ = [
        ("demInitialRepeat", C.c_double), ("maxTime", C.c_double),
        ("demSolve", C.c_int), ("lbSolve", C.c_int), ("freeSurface", C.c_int), ("forceField", C.c_int),
        ("unitLength", C.c_double), ("unitTime", C.c_double), ("unitDensity", C.c_double),
        ("lbSize", C.c_double * 3), ("initVelocity", C.c_double * 3), ("lbF", C.c_double * 3),
        ("boundary", C.c_int * 6), ("slipCoefficient", C.c_double), ("fluidHeight", C.c_double),
        ("initVisc", C.c_double),
        ("rheologyModel", C.c_int), ("plasticVisc", C.c_double), ("yieldStress", C.c_double),
        ("frictionCoefFluid", C.c_double), ("voellmyConstant", C.c_double), ("turbulenceSolve", C.c_int),
        ("turbConst", C.c_double),
        ("density", C.c_double), ("objectDensity", C.c_double), ("bodyDensity", C.c_double),
        ("linearStiff", C.c_double), ("frictionCoefPart", C.c_double),
        ("lubNormal", C.c_double), ("lubTangential", C.c_double), ("lubRotation", C.c_double),
        ("lubCutoff", C.c_double), ("kinVisc", C.c_double),
        ("electroForce", C.c_double), ("debyeLength", C.c_double),
        ("objWaitTime", C.c_int),
        ("springDamp", C.c_double), ("epsilon", C.c_double), ("k13", C.c_double), ("nonLinType", C.c_int),
        ("precision", C.c_int), ("device", C.c_int),
    ]


class Info(C.Structure):
    _fields_ = [
        ("size", C.c_int * 3), ("totNodes", C.c_int64),
        ("nParticles", C.c_int), ("nObjects", C.c_int), ("nWalls", C.c_int),
        ("nActive", C.c_int), ("nFluid", C.c_int), ("nInterface", C.c_int), ("nParticleNodes", C.c_int),
        ("nObjectNodes", C.c_int),
        ("lbTime", C.c_uint), ("demTimeStep", C.c_uint), ("demTime", C.c_double), ("realTime", C.c_double),
        ("neighborListBuilds", C.c_int), ("newNeighborList", C.c_int),
        ("totalMassInit", C.c_double), ("lastExtraMass", C.c_double), ("lastSurplus", C.c_double),
        ("precision", C.c_int), ("nBodies", C.c_int),
    ]

    def as_dict(self) -> Dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "size" else v
        return d


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

RUN_QUEUED, RUN_ACTIVE, RUN_DONE, RUN_FAILED, RUN_CANCELLED = 0, 1, 2, 3, 4


class RunDesc(C.Structure):
    _fields_ = [
        ("params", Params), ("nParticles", C.c_int), ("size", _ip), ("r", _dp), ("x0", _dp),
        ("nObjects", C.c_int), ("objR", _dp), ("objX0", _dp), ("objX1", _dp), ("nsteps", C.c_int),
    ]


class RunResult(C.Structure):
    _fields_ = [
        ("state", C.c_int), ("error", C.c_int), ("stepsDone", C.c_int), ("nObjects", C.c_int),
        ("tSubmit", C.c_double), ("tStart", C.c_double), ("tDone", C.c_double),
        ("objects", C.c_double * (2 * 38)), ("totalMass", C.c_double), ("nActive", C.c_int), ("tileZ", C.c_int),
        ("nodeSteps", C.c_int64), ("deviceBytes", C.c_size_t), ("message", C.c_char * 128),
    ]

    def as_dict(self) -> Dict:
        no = self.nObjects
        return dict(state=self.state, error=self.error, stepsDone=self.stepsDone, nObjects=no,
                    tSubmit=self.tSubmit, tStart=self.tStart, tDone=self.tDone,
                    objects=np.array(self.objects[:no * 38], dtype=np.float64).reshape(no, 38),
                    totalMass=self.totalMass, nActive=self.nActive, tileZ=self.tileZ, nodeSteps=self.nodeSteps,
                    deviceBytes=self.deviceBytes, message=self.message.decode(errors="replace"))


class Tuning(C.Structure):
    _fields_ = [("tile", C.c_int * 3), ("tuned", C.c_int), ("nCandidates", C.c_int), ("candidateZ", C.c_int * 4),
                ("candidateMs", C.c_double * 4), ("pairCaptured", C.c_int)]

    def as_dict(self) -> Dict:
        n = self.nCandidates
        return dict(tile=list(self.tile), tuned=bool(self.tuned),
                    candidates={int(self.candidateZ[i]): float(self.candidateMs[i]) for i in range(n)},
                    pair_captured=bool(self.pairCaptured))


class PoolStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int), ("queued", C.c_int), ("active", C.c_int), ("done", C.c_int),
        ("failed", C.c_int), ("cancelled", C.c_int), ("activePeak", C.c_int), ("batchBuilds", C.c_int),
        ("batchedPeak", C.c_int),
        ("budget", C.c_size_t), ("bytesInUse", C.c_size_t), ("bytesPeak", C.c_size_t),
    ]

    def as_dict(self) -> Dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class LbdemError(RuntimeError):
    pass


class Library:
    """Loads a shared library that exports the lbdem C ABI."""

    _cache: Dict[str, "Library"] = {}

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise LbdemError(f"lbdem library not found: {path} (build it with __graft_entry__.build())")
        self.path = path
        lib = C.CDLL(path, mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2))
        vp = C.c_void_p
        sig = {
            "lbdem_create": ([C.POINTER(Params), C.POINTER(vp)], C.c_int),
            "lbdem_destroy": ([vp], C.c_int),
            "lbdem_last_error": ([vp], C.c_char_p),
            "lbdem_set_particles": ([vp, C.c_int, _ip, _dp, _dp], C.c_int),
            "lbdem_set_objects": ([vp, C.c_int, _dp, _dp, _dp], C.c_int),
            "lbdem_set_restart_fluid": ([vp, _ip, C.c_void_p, C.c_int, _dp], C.c_int),
            "lbdem_store_velocity": ([vp], C.c_int),
            "lbdem_get_velocity": ([vp, C.c_void_p, C.c_void_p], C.c_int),
            "lbdem_init": ([vp], C.c_int),
            "lbdem_dem_step": ([vp], C.c_int),
            "lbdem_coupling_step": ([vp], C.c_int),
            "lbdem_lb_step": ([vp], C.c_int),
            "lbdem_fs_step": ([vp], C.c_int),
            "lbdem_cycle": ([vp, C.c_int], C.c_int),
            "lbdem_synchronize": ([vp], C.c_int),
            "lbdem_get_info": ([vp, C.POINTER(Info)], C.c_int),
            "lbdem_get_types": ([vp, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "lbdem_get_fields": ([vp, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "lbdem_get_fields_noalloc": ([vp, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "lbdem_get_particles": ([vp, _dp], C.c_int),
            "lbdem_get_objects": ([vp, _dp], C.c_int),
            "lbdem_set_bodies": ([vp, C.c_int, _dp, _dp, _dp, _dp, _dp], C.c_int),
            "lbdem_get_bodies": ([vp, _dp], C.c_int),
            "lbdem_set_stress": ([vp, C.c_int], C.c_int),
            "lbdem_get_stress": ([vp, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "lbdem_get_particle_stress": ([vp, _dp], C.c_int),
            "lbdem_get_walls": ([vp, _dp], C.c_int),
            "lbdem_get_neighbors": ([vp, C.c_int, _ip, _ip], C.c_int),
            "lbdem_total_mass": ([vp, _dp], C.c_int),
            "lbdem_profile_enable": ([vp, C.c_int], C.c_int),
            "lbdem_profile_read": ([vp, _dp, _ip], C.c_int),
            "lbdem_estimate_bytes": ([C.POINTER(Params), C.c_int, _dp, C.c_int, _dp, C.POINTER(C.c_size_t)], C.c_int),
            "lbdem_device_bytes": ([vp, C.POINTER(C.c_size_t)], C.c_int),
            "lbdem_use_graphs": ([vp, C.c_int], C.c_int),
            "lbdem_set_tile": ([vp, C.c_int], C.c_int),
            "lbdem_get_tuning": ([vp, C.POINTER(Tuning)], C.c_int),
            "lbdem_batch_create": ([C.POINTER(vp), C.c_int, C.POINTER(vp)], C.c_int),
            "lbdem_batch_cycle": ([vp, C.c_int], C.c_int),
            "lbdem_batch_synchronize": ([vp], C.c_int),
            "lbdem_batch_last_error": ([vp], C.c_char_p),
            "lbdem_batch_destroy": ([vp], C.c_int),
            "lbdem_batch_set_tile": ([vp, C.c_int], C.c_int),
            "lbdem_batch_get_tile": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
            "lbdem_batch_profile_enable": ([vp, C.c_int], C.c_int),
            "lbdem_batch_profile_read": ([vp, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
            "lbdem_pool_create": ([C.c_size_t, C.c_int, C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
            "lbdem_pool_submit": ([vp, C.POINTER(RunDesc), _ip], C.c_int),
            "lbdem_pool_poll": ([vp, C.c_int, C.POINTER(RunResult)], C.c_int),
            "lbdem_pool_wait": ([vp, C.c_int, C.c_double, C.POINTER(RunResult)], C.c_int),
            "lbdem_pool_drain": ([vp, C.c_double], C.c_int),
            "lbdem_pool_cancel": ([vp, C.c_int], C.c_int),
            "lbdem_pool_get_stats": ([vp, C.POINTER(PoolStats)], C.c_int),
            "lbdem_pool_set_batching": ([vp, C.c_int], C.c_int),
            "lbdem_pool_set_batch_tile": ([vp, C.c_int], C.c_int),
            "lbdem_pool_clock": ([vp], C.c_double),
            "lbdem_pool_last_error": ([vp], C.c_char_p),
            "lbdem_pool_destroy": ([vp], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        self.lib = lib

    @classmethod
    def load(cls, path: Optional[str] = None) -> "Library":
        path = os.path.abspath(path or LIB_PATH)
        if path not in cls._cache:
            cls._cache[path] = Library(path)
        return cls._cache[path]


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Simulation:
    """One lbdem context: configuration, particles and objects, then steps and downloads."""

    def __init__(self, lib: Library, params: Params, particles: "ParticleSet", objects: "ObjectSet"):
        self.lib = lib
        self.ctx = C.c_void_p()
        self.params = params
        rc = lib.lib.lbdem_create(C.byref(params), C.byref(self.ctx))
        self._check(rc, "lbdem_create")
        n = len(particles.r)
        size = np.ascontiguousarray(particles.size, dtype=np.int32)
        r = np.ascontiguousarray(particles.r, dtype=np.float64)
        x0 = np.ascontiguousarray(particles.x0, dtype=np.float64).reshape(-1)
        self._check(lib.lib.lbdem_set_particles(self.ctx, n, _ptr(size, C.c_int), _ptr(r, C.c_double),
                                                _ptr(x0, C.c_double)), "lbdem_set_particles")
        m = len(objects.r)
        orr = np.ascontiguousarray(objects.r, dtype=np.float64)
        ox0 = np.ascontiguousarray(objects.x0, dtype=np.float64).reshape(-1)
        ox1 = np.ascontiguousarray(objects.x1, dtype=np.float64).reshape(-1)
        self._check(lib.lib.lbdem_set_objects(self.ctx, m, _ptr(orr, C.c_double), _ptr(ox0, C.c_double),
                                              _ptr(ox1, C.c_double)), "lbdem_set_objects")
        bs = objects.bodies
        if bs is not None and len(bs.r):
            arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (bs.r, bs.x0, bs.x1, bs.k, bs.L0)]
            self._check(lib.lib.lbdem_set_bodies(self.ctx, len(bs.r), *[_ptr(a, C.c_double) for a in arrs]),
                        "lbdem_set_bodies")

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.lib.lbdem_last_error(self.ctx) if self.ctx else b""
            raise LbdemError(f"{what} failed with code {rc}: {msg.decode(errors='replace') if msg else ''}")

    def close(self):
        if self.ctx:
            self.lib.lib.lbdem_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_restart(self, size, types: np.ndarray, records: np.ndarray):
        """The contents of a fluidRecycle file (read_recycle) for lbdem_init to restart from."""
        sz = np.ascontiguousarray(size, dtype=np.int32)
        t = np.ascontiguousarray(types, dtype=np.uint8)
        rec = np.ascontiguousarray(records, dtype=np.float64).reshape(-1, RESTART_REC)
        self._check(self.lib.lib.lbdem_set_restart_fluid(self.ctx, _ptr(sz, C.c_int), t.ctypes.data, len(rec),
                                                         _ptr(rec, C.c_double)), "lbdem_set_restart_fluid")

    def init(self):
        self._check(self.lib.lib.lbdem_init(self.ctx), "lbdem_init")

    def set_stress(self, enable: bool = True):
        """Compute the stress diagnostics (before init)."""
        self._check(self.lib.lib.lbdem_set_stress(self.ctx, 1 if enable else 0), "lbdem_set_stress")

    def stress(self):
        """(global 5 x 3 x 3: totalStress, hydroStress, lubStress, contactStress, electroStress; STotal
        of every particle, n x 3 x 3; disp, n x 3) of the last DEM step."""
        n = self.info()["nParticles"]
        g = np.zeros(45, np.float64)
        st = np.zeros(max(n, 1) * 9, np.float64)
        d = np.zeros(max(n, 1) * 3, np.float64)
        self._check(self.lib.lib.lbdem_get_stress(self.ctx, g.ctypes.data, st.ctypes.data, d.ctypes.data),
                    "lbdem_get_stress")
        return g.reshape(5, 3, 3), st[:9 * n].reshape(n, 3, 3), d[:3 * n].reshape(n, 3)

    def particle_stress(self) -> np.ndarray:
        """n x 44 records of lbdem_get_particle_stress (STotal, SElas, SVisc, SVisc_h, overlap, disp,
        coord, lub_coord) of the last DEM step."""
        n = self.info()["nParticles"]
        rec = np.zeros((max(n, 1), 44), np.float64)
        self._check(self.lib.lib.lbdem_get_particle_stress(self.ctx, _ptr(rec, C.c_double)),
                    "lbdem_get_particle_stress")
        return rec[:n]

    def store_velocity(self):
        """The next LB step keeps every active node's velocity (read with velocity())."""
        self._check(self.lib.lib.lbdem_store_velocity(self.ctx), "lbdem_store_velocity")

    def velocity(self):
        """(u as N x 3, viscosity as N) of every active node, zero elsewhere."""
        N = self.info()["totNodes"]
        u = np.zeros(3 * N, np.float64)
        v = np.zeros(N, np.float64)
        self._check(self.lib.lib.lbdem_get_velocity(self.ctx, u.ctypes.data, v.ctypes.data), "lbdem_get_velocity")
        return u.reshape(N, 3), v

    def dem_step(self):
        self._check(self.lib.lib.lbdem_dem_step(self.ctx), "lbdem_dem_step")

    def coupling_step(self):
        self._check(self.lib.lib.lbdem_coupling_step(self.ctx), "lbdem_coupling_step")

    def lb_step(self):
        self._check(self.lib.lib.lbdem_lb_step(self.ctx), "lbdem_lb_step")

    def fs_step(self):
        self._check(self.lib.lib.lbdem_fs_step(self.ctx), "lbdem_fs_step")

    def cycle(self, n: int = 1):
        self._check(self.lib.lib.lbdem_cycle(self.ctx, int(n)), "lbdem_cycle")

    def synchronize(self):
        self._check(self.lib.lib.lbdem_synchronize(self.ctx), "lbdem_synchronize")

    def info(self) -> Dict:
        inf = Info()
        self._check(self.lib.lib.lbdem_get_info(self.ctx, C.byref(inf)), "lbdem_get_info")
        return inf.as_dict()

    def types(self):
        N = self.info()["totNodes"]
        t = np.zeros(N, np.uint8)
        p = np.zeros(N, np.uint8)
        s = np.zeros(N, np.int32)
        self._check(self.lib.lib.lbdem_get_types(self.ctx, t.ctypes.data, p.ctypes.data, s.ctypes.data),
                    "lbdem_get_types")
        return t, p, s

    def fields(self, want_f: bool = True, noalloc: bool = False):
        """Populations (19 x N, or None), mass and density fields; ``noalloc`` goes through
        lbdem_get_fields_noalloc (the library's static scratch; boxes of at most 2^21 nodes)."""
        N = self.info()["totNodes"]
        f = np.zeros(Q * N, np.float64) if want_f else None
        m = np.zeros(N, np.float64)
        n = np.zeros(N, np.float64)
        fn = self.lib.lib.lbdem_get_fields_noalloc if noalloc else self.lib.lib.lbdem_get_fields
        self._check(fn(self.ctx, f.ctypes.data if want_f else None, m.ctypes.data, n.ctypes.data),
                    "lbdem_get_fields_noalloc" if noalloc else "lbdem_get_fields")
        return (f.reshape(Q, N) if want_f else None), m, n

    def particles(self) -> np.ndarray:
        n = self.info()["nParticles"]
        rec = np.zeros((max(n, 1), P["REC"]), np.float64)
        self._check(self.lib.lib.lbdem_get_particles(self.ctx, _ptr(rec, C.c_double)), "lbdem_get_particles")
        return rec[:n]

    def objects(self) -> np.ndarray:
        n = self.info()["nObjects"]
        rec = np.zeros((max(n, 1), O["REC"]), np.float64)
        self._check(self.lib.lib.lbdem_get_objects(self.ctx, _ptr(rec, C.c_double)), "lbdem_get_objects")
        return rec[:n]

    def bodies(self) -> np.ndarray:
        n = self.info()["nBodies"]
        rec = np.zeros((max(n, 1), B["REC"]), np.float64)
        self._check(self.lib.lib.lbdem_get_bodies(self.ctx, _ptr(rec, C.c_double)), "lbdem_get_bodies")
        return rec[:n]

    def walls(self) -> np.ndarray:
        n = self.info()["nWalls"]
        rec = np.zeros((max(n, 1), W["REC"]), np.float64)
        self._check(self.lib.lib.lbdem_get_walls(self.ctx, _ptr(rec, C.c_double)), "lbdem_get_walls")
        return rec[:n]

    def neighbors(self, cap: int = 64):
        n = self.info()["nParticles"]
        cnt = np.zeros(max(n, 1), np.int32)
        lst = np.full((max(n, 1), cap), -1, np.int32)
        self._check(self.lib.lib.lbdem_get_neighbors(self.ctx, cap, _ptr(cnt, C.c_int), _ptr(lst, C.c_int)),
                    "lbdem_get_neighbors")
        return cnt[:n], lst[:n]

    def profile(self, enable: bool = True):
        self._check(self.lib.lib.lbdem_profile_enable(self.ctx, 1 if enable else 0), "lbdem_profile_enable")

    def profile_read(self):
        ms = C.c_double()
        n = C.c_int()
        self._check(self.lib.lib.lbdem_profile_read(self.ctx, C.byref(ms), C.byref(n)), "lbdem_profile_read")
        return ms.value, n.value

    def total_mass(self) -> float:
        v = C.c_double()
        self._check(self.lib.lib.lbdem_total_mass(self.ctx, C.byref(v)), "lbdem_total_mass")
        return v.value

    def use_graphs(self, enable: bool):
        self._check(self.lib.lib.lbdem_use_graphs(self.ctx, 1 if enable else 0), "lbdem_use_graphs")

    def set_tile(self, tile_z: int):
        """0: tune the fused kernel's z-walk length in the first LB steps; 1, 2, 4, 8: fix it."""
        self._check(self.lib.lib.lbdem_set_tile(self.ctx, int(tile_z)), "lbdem_set_tile")

    def tuning(self) -> Dict:
        t = Tuning()
        self._check(self.lib.lib.lbdem_get_tuning(self.ctx, C.byref(t)), "lbdem_get_tuning")
        return t.as_dict()

    def device_bytes(self) -> int:
        v = C.c_size_t()
        self._check(self.lib.lib.lbdem_device_bytes(self.ctx, C.byref(v)), "lbdem_device_bytes")
        return v.value


def estimate_bytes(lib: Library, params: Params, particles: "ParticleSet", objects: "ObjectSet") -> int:
    """Device bytes lbdem_init will take for this configuration (host-side computation)."""
    r = np.ascontiguousarray(particles.r, dtype=np.float64)
    orr = np.ascontiguousarray(objects.r, dtype=np.float64)
    v = C.c_size_t()
    rc = lib.lib.lbdem_estimate_bytes(C.byref(params), len(r), _ptr(r, C.c_double), len(orr), _ptr(orr, C.c_double),
                                      C.byref(v))
    if rc != 0:
        raise LbdemError(f"lbdem_estimate_bytes failed with code {rc}")
    return v.value


class Batch:
    """Contexts stepped together: each kernel of a step launched once for all of them."""

    def __init__(self, lib: Library, sims: Sequence["Simulation"]):
        self.lib = lib
        self.sims = list(sims)
        self.batch = C.c_void_p()
        arr = (C.c_void_p * len(self.sims))(*[sm.ctx.value for sm in self.sims])
        rc = lib.lib.lbdem_batch_create(arr, len(self.sims), C.byref(self.batch))
        if rc != 0:
            msgs = [sm.lib.lib.lbdem_last_error(sm.ctx).decode(errors="replace") for sm in self.sims]
            raise LbdemError(f"lbdem_batch_create failed with code {rc}: {[m for m in msgs if m]}")

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.lib.lbdem_batch_last_error(self.batch) if self.batch else b""
            raise LbdemError(f"{what} failed with code {rc}: {msg.decode(errors='replace') if msg else ''}")

    def cycle(self, n: int):
        self._check(self.lib.lib.lbdem_batch_cycle(self.batch, int(n)), "lbdem_batch_cycle")

    def synchronize(self):
        self._check(self.lib.lib.lbdem_batch_synchronize(self.batch), "lbdem_batch_synchronize")

    def set_tile(self, tile_z: int):
        """z-walk for every member (1, 2, 4, 8); 0: tune it over the next step pairs."""
        self._check(self.lib.lib.lbdem_batch_set_tile(self.batch, int(tile_z)), "lbdem_batch_set_tile")

    def tile(self) -> Dict:
        tz, tuned = C.c_int(), C.c_int()
        self._check(self.lib.lib.lbdem_batch_get_tile(self.batch, C.byref(tz), C.byref(tuned)), "lbdem_batch_get_tile")
        return dict(tileZ=tz.value, tuned=bool(tuned.value))

    def profile(self, enable: bool = True):
        """Time every batch launch with HIP events (pairs run launch by launch, no graph)."""
        self._check(self.lib.lib.lbdem_batch_profile_enable(self.batch, 1 if enable else 0),
                    "lbdem_batch_profile_enable")

    def profile_read(self) -> List[Dict]:
        """[{kernel, ms, launches}] since the last enable, in order of first launch."""
        out, i = [], 0
        while True:
            name, ms, n, cnt = C.c_char_p(), C.c_double(), C.c_int(), C.c_int()
            self._check(self.lib.lib.lbdem_batch_profile_read(self.batch, i, C.byref(name), C.byref(ms), C.byref(n),
                                                              C.byref(cnt)), "lbdem_batch_profile_read")
            if i >= cnt.value:
                return out
            out.append(dict(kernel=name.value.decode(), ms=ms.value, launches=n.value))
            i += 1

    def close(self):
        if self.batch:
            self.lib.lib.lbdem_batch_destroy(self.batch)
            self.batch = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pool:
    """Run pool of the device library: independent runs admitted against an HBM budget, each on
    its own stream, stepped concurrently; results are the impactor records, total mass and
    active-node count after the last step."""

    def __init__(self, lib: Library, budget_bytes: int = 0, max_concurrent: int = 64, steps_per_chunk: int = 32,
                 device: int = 0, batching: bool = True, batch_tile: int = 0):
        self.lib = lib
        self.pool = C.c_void_p()
        rc = lib.lib.lbdem_pool_create(int(budget_bytes), int(max_concurrent), int(steps_per_chunk), int(device),
                                       C.byref(self.pool))
        if rc != 0:
            raise LbdemError(f"lbdem_pool_create failed with code {rc}")
        self._check(lib.lib.lbdem_pool_set_batching(self.pool, 1 if batching else 0), "lbdem_pool_set_batching")
        self._check(lib.lib.lbdem_pool_set_batch_tile(self.pool, int(batch_tile)), "lbdem_pool_set_batch_tile")

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.lib.lbdem_pool_last_error(self.pool) if self.pool else b""
            raise LbdemError(f"{what} failed with code {rc}: {msg.decode(errors='replace') if msg else ''}")

    def submit(self, params: Params, particles: "ParticleSet", objects: "ObjectSet", nsteps: int) -> int:
        size = np.ascontiguousarray(particles.size, dtype=np.int32)
        r = np.ascontiguousarray(particles.r, dtype=np.float64)
        x0 = np.ascontiguousarray(particles.x0, dtype=np.float64).reshape(-1)
        orr = np.ascontiguousarray(objects.r, dtype=np.float64)
        ox0 = np.ascontiguousarray(objects.x0, dtype=np.float64).reshape(-1)
        ox1 = np.ascontiguousarray(objects.x1, dtype=np.float64).reshape(-1)
        d = RunDesc()
        d.params = params
        d.nParticles = len(r)
        d.size, d.r, d.x0 = _ptr(size, C.c_int), _ptr(r, C.c_double), _ptr(x0, C.c_double)
        d.nObjects = len(orr)
        d.objR, d.objX0, d.objX1 = _ptr(orr, C.c_double), _ptr(ox0, C.c_double), _ptr(ox1, C.c_double)
        d.nsteps = int(nsteps)
        rid = C.c_int()
        self._check(self.lib.lib.lbdem_pool_submit(self.pool, C.byref(d), C.byref(rid)), "lbdem_pool_submit")
        return rid.value

    def poll(self, rid: int) -> Dict:
        res = RunResult()
        self._check(self.lib.lib.lbdem_pool_poll(self.pool, rid, C.byref(res)), "lbdem_pool_poll")
        return res.as_dict()

    def wait(self, rid: int, timeout: float = 600.0) -> Dict:
        res = RunResult()
        self._check(self.lib.lib.lbdem_pool_wait(self.pool, rid, float(timeout), C.byref(res)), "lbdem_pool_wait")
        return res.as_dict()

    def drain(self, timeout: float = 600.0):
        self._check(self.lib.lib.lbdem_pool_drain(self.pool, float(timeout)), "lbdem_pool_drain")

    def cancel(self, rid: int):
        self._check(self.lib.lib.lbdem_pool_cancel(self.pool, rid), "lbdem_pool_cancel")

    def stats(self) -> Dict:
        st = PoolStats()
        self._check(self.lib.lib.lbdem_pool_get_stats(self.pool, C.byref(st)), "lbdem_pool_get_stats")
        return st.as_dict()

    def clock(self) -> float:
        return self.lib.lib.lbdem_pool_clock(self.pool)

    def close(self):
        if self.pool:
            self.lib.lib.lbdem_pool_destroy(self.pool)
            self.pool = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------------------------
# getpot configuration (reference keys and defaults)
# ----------------------------------------------------------------------------------------------
DEFAULTS: Dict[str, object] = {
    # main.cpp:131-175
    "demInitialRepeat": 0.0, "lbmInitialRepeat": 0, "maxTime": 0.0, "maximumTimeSteps": 0,
    "screenExpTime": 0.0, "fluidExpTime": 0.0, "fluidLagrangianExpTime": 0.0, "partExpTime": 0.0,
    "fluidRecycleExpTime": 0.0, "partRecycleExpTime": 0.0,
    "demSolve": 0, "lbSolve": 0, "freeSurfaceSolve": 0, "forceFieldSolve": 0,
    # LB.cpp:137-336
    "lbRestart": 0, "lbTopography": 0, "unitLength": 1.0, "unitTime": 1.0, "unitDensity": 1.0,
    "lbSizeX": 0.0, "lbSizeY": 0.0, "lbSizeZ": 0.0, "rheologyModel": "none", "initVisc": 1.0,
    "plasticVisc": 0.0, "yieldStress": 0.0, "frictionCoefFluid": 0.0, "voellmyConstant": 1.0,
    "turbulenceSolve": 0, "turbConst": 0.0,
    "initVelocityX": 0.0, "initVelocityY": 0.0, "initVelocityZ": 0.0,
    "lbFX": 0.0, "lbFY": 0.0, "lbFZ": 0.0,
    "boundary0": 1, "boundary1": 1, "boundary2": 1, "boundary3": 1, "boundary4": 1, "boundary5": 1,
    "slipCoefficient": 0.0, "fluidHeight": 0.0,
    # DEM.cpp:19-75
    "density": 0.0, "objectDensity": 0.0, "bodyDensity": 0.0, "linearStiff": 0.0, "frictionCoefPart": 0.0,
    "lubNormal": 0.0, "lubTangential": 0.0, "lubRotation": 0.0, "lubCutoff": 0.0, "kinVisc": 0.0,
    "electroForce": 0.0, "debyeLength": 0.0, "objWaitTime": 0, "intervalTime": 0, "springDamp": 0.0,
    "epsilon": 0.0, "nonLinType": 0, "k13": 0.0, "scale": 1.0,
}

_LINE = re.compile(r"^\s*([A-Za-z_][A-Za-z0-9_]*)\s*=\s*(.*?)\s*$")


def parse_getpot(text: str) -> Dict[str, str]:
    """Variables of a getpot file: ``name = value``, ``#`` starts a comment, quotes allowed."""
    out: Dict[str, str] = {}
    for raw in text.splitlines():
        line = raw.split("#", 1)[0]
        m = _LINE.match(line)
        if not m:
            continue
        val = m.group(2).strip()
        if len(val) >= 2 and val[0] == val[-1] and val[0] in "'\"":
            val = val[1:-1]
        out[m.group(1)] = val
    return out


def _num(v: str, like):
    """Value of a getpot variable, typed like the reference default (strings stay strings)."""
    if isinstance(like, str):
        return v
    s = v.strip().lower()
    if s in ("true", "false"):
        return 1 if s == "true" else 0
    try:
        x = float(s)
    except ValueError:
        return v
    if isinstance(like, int) and not isinstance(like, bool):
        return int(x)
    return x


def config_values(text: str, overrides: Optional[Dict[str, object]] = None) -> Dict[str, object]:
    vals = dict(DEFAULTS)
    for k, v in parse_getpot(text).items():
        vals[k] = _num(v, DEFAULTS.get(k, 0.0))
    if overrides:
        vals.update(overrides)
    return vals


RHEOLOGY = {"NEWTONIAN": RHEO_NEWTONIAN, "BINGHAM": RHEO_BINGHAM, "FRICTIONAL": RHEO_FRICTIONAL,
            "VOELLMY": RHEO_VOELLMY}


def params_from_values(v: Dict[str, object], precision: int = PREC_FP64, device: int = 0) -> Params:
    """lbdem_params from getpot values. The rheology keys are read as LB.cpp:197-245 reads them: an
    unknown model name keeps NEWTONIAN; a BINGHAM run reads no initVisc and keeps the default
    (1.0 - 0.5) / 3.0, which LB.cpp:320 then scales like a physical viscosity."""
    p = Params()
    model = RHEOLOGY.get(str(v.get("rheologyModel", "none")), RHEO_NEWTONIAN)
    p.rheologyModel = model
    p.initVisc = (1.0 - 0.5) / 3.0 if model == RHEO_BINGHAM else float(v["initVisc"])
    if model == RHEO_BINGHAM:
        p.plasticVisc, p.yieldStress = float(v.get("plasticVisc", 0.0)), float(v.get("yieldStress", 0.0))
    if model in (RHEO_FRICTIONAL, RHEO_VOELLMY):
        p.frictionCoefFluid = float(v.get("frictionCoefFluid", 0.0))
    if model == RHEO_VOELLMY:
        p.voellmyConstant = float(v.get("voellmyConstant", 1.0))
    p.turbulenceSolve = int(v.get("turbulenceSolve", 0))
    p.turbConst = float(v.get("turbConst", 0.0))
    p.demInitialRepeat = float(v["demInitialRepeat"])
    p.maxTime = float(v["maxTime"])
    p.demSolve, p.lbSolve = int(v["demSolve"]), int(v["lbSolve"])
    p.freeSurface, p.forceField = int(v["freeSurfaceSolve"]), int(v["forceFieldSolve"])
    p.unitLength, p.unitTime, p.unitDensity = float(v["unitLength"]), float(v["unitTime"]), float(v["unitDensity"])
    p.lbSize[:] = [float(v["lbSizeX"]), float(v["lbSizeY"]), float(v["lbSizeZ"])]
    p.initVelocity[:] = [float(v["initVelocityX"]), float(v["initVelocityY"]), float(v["initVelocityZ"])]
    p.lbF[:] = [float(v["lbFX"]), float(v["lbFY"]), float(v["lbFZ"])]
    p.boundary[:] = [int(v[f"boundary{i}"]) for i in range(6)]
    p.fluidHeight = float(v["fluidHeight"])
    p.slipCoefficient = float(v["slipCoefficient"])
    for k in ("density", "objectDensity", "bodyDensity", "linearStiff", "frictionCoefPart", "lubNormal",
              "lubTangential", "lubRotation", "lubCutoff", "kinVisc", "electroForce", "debyeLength"):
        setattr(p, k, float(v[k]))
    p.objWaitTime = int(v["objWaitTime"])
    p.springDamp, p.epsilon, p.k13 = float(v["springDamp"]), float(v["epsilon"]), float(v["k13"])
    p.nonLinType = int(v["nonLinType"])
    p.precision = int(precision)
    p.device = int(device)
    return p


def write_getpot(path: str, values: Dict[str, object]):
    with open(path, "w") as fh:
        for k, v in values.items():
            if isinstance(v, (float, np.floating)):
                fh.write(f"{k} = {float(v)!r}\n")
            else:
                fh.write(f"{k} = {v}\n")


# ----------------------------------------------------------------------------------------------
# particles.dat / objects.dat
# ----------------------------------------------------------------------------------------------
@dataclass
class ParticleSet:
    size: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    r: np.ndarray = field(default_factory=lambda: np.zeros(0))
    x0: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    # velocity columns as read (particles.dat columns 6-11); the solver starts every sphere at
    # rest, as particle::initialize resets x1 and w1 (elmt.cpp:28-31, called from DEM.cpp:232)
    x1: Optional[np.ndarray] = None
    w1: Optional[np.ndarray] = None


@dataclass
class BodySet:
    """input/body.dat rows (DEM.cpp:168-196): radius, centre (scaled), velocity, spring stiffness k
    and rest length L0."""
    r: np.ndarray = field(default_factory=lambda: np.zeros(0))
    x0: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    x1: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    L0: np.ndarray = field(default_factory=lambda: np.zeros(0))


@dataclass
class ObjectSet:
    r: np.ndarray = field(default_factory=lambda: np.zeros(0))
    x0: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    x1: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    bodies: Optional[BodySet] = None  # the spring-connected bodies of input/body.dat, if any


def _tokens(path: str) -> List[str]:
    with open(path) as fh:
        return fh.read().split()


def read_particles(path: str, scale: float = 1.0) -> ParticleSet:
    """input/particles.dat as DEM::discreteElementGet reads it (DEM.cpp:81-134)."""
    tok = _tokens(path)
    n = int(tok[0])
    rows = np.array([float(x) for x in tok[1:1 + 20 * n]], dtype=np.float64).reshape(n, 20) if n else np.zeros((0, 20))
    ps = ParticleSet()
    ps.size = rows[:, 1].astype(np.int32)
    ps.r = rows[:, 2] * scale
    ps.x0 = rows[:, 3:6] * scale - 0.05
    if n and (np.any(rows[:, 6:12] != 0.0)):
        ps.x1 = rows[:, 6:9].copy()
        ps.w1 = rows[:, 9:12].copy()
    return ps


def read_objects(path: str, scale: float = 1.0) -> ObjectSet:
    """input/objects.dat as DEM::discreteElementGet reads it (DEM.cpp:136-165)."""
    tok = _tokens(path)
    n = int(tok[0])
    rows = np.array([float(x) for x in tok[1:1 + 8 * n]], dtype=np.float64).reshape(n, 8) if n else np.zeros((0, 8))
    os_ = ObjectSet()
    os_.r = rows[:, 1] * scale
    os_.x0 = rows[:, 2:5] * scale
    os_.x1 = rows[:, 5:8].copy()
    return os_


def write_particles(path: str, r: Sequence[float], centers: np.ndarray, size: Optional[Sequence[int]] = None,
                    x1: Optional[np.ndarray] = None, w1: Optional[np.ndarray] = None):
    """Rows in particles.dat format; ``centers`` are written as given (the reader subtracts 0.05)."""
    centers = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    n = len(r)
    x1 = np.zeros((n, 3)) if x1 is None else np.asarray(x1, dtype=np.float64).reshape(-1, 3)
    w1 = np.zeros((n, 3)) if w1 is None else np.asarray(w1, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write(f"{n}\n")
        for i in range(n):
            s = 1 if size is None else int(size[i])
            x, y, z = (float(c) for c in centers[i])
            v = " ".join(repr(float(c)) for c in (*x1[i], *w1[i]))
            fh.write(f"{i} {s} {float(r[i])!r} {x!r} {y!r} {z!r} {v} 1 0 0 0 0 0 0 0\n")


def write_objects(path: str, r: Sequence[float], x0: np.ndarray, x1: np.ndarray):
    x0 = np.asarray(x0, dtype=np.float64).reshape(-1, 3)
    x1 = np.asarray(x1, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write(f"{len(r)}\n")
        for i in range(len(r)):
            a = [float(c) for c in (*x0[i], *x1[i])]
            fh.write(f"{i} {float(r[i])!r} " + " ".join(repr(c) for c in a) + "\n")


RESTART_REC = 25  # include/lbdem.h LBDEM_RESTART_REC: n - 1, u (3), mass - 1, visc, f (19)


def read_recycle(path: str):
    """A fluidRecycle file (IO.cpp:973-997): (lattice size, type of every node, records as
    n_records x 25), parsed as the reference's restart reads it (LB.cpp:352-367, 994-1009,
    1098-1138): whitespace-separated numbers."""
    with open(path) as fh:
        tok = fh.read().split()
    size = [int(t) for t in tok[:3]]
    n = size[0] * size[1] * size[2]
    types = np.array([int(t) for t in tok[3:3 + n]], np.uint8)
    vals = np.array([float(t) for t in tok[3 + n:]], np.float64)
    m = len(vals) // RESTART_REC
    return size, types, vals[:m * RESTART_REC].reshape(m, RESTART_REC)


def read_bodies(path: str, scale: float = 1.0) -> BodySet:
    """input/body.dat as DEM::discreteElementGet reads it (DEM.cpp:168-196): a count, then index, r,
    x0 (times scale), x1, k, L0 per body."""
    tok = _tokens(path)
    n = int(tok[0])
    rows = np.array([float(x) for x in tok[1:1 + 10 * n]], dtype=np.float64).reshape(n, 10) if n else np.zeros((0, 10))
    return BodySet(r=rows[:, 1].copy(), x0=rows[:, 2:5] * scale, x1=rows[:, 5:8].copy(), k=rows[:, 8].copy(),
                   L0=rows[:, 9].copy())


def write_bodies(path: str, r: Sequence[float], x0: np.ndarray, x1: np.ndarray, k: Sequence[float],
                 L0: Sequence[float]):
    with open(path, "w") as fh:
        fh.write(f"{len(r)}\n")
        for i in range(len(r)):
            fh.write(f"{i} {r[i]!r} {x0[i][0]!r} {x0[i][1]!r} {x0[i][2]!r} {x1[i][0]!r} {x1[i][1]!r} {x1[i][2]!r} "
                     f"{k[i]!r} {L0[i]!r}\n")


def load_case(case_dir: str, cfg_name: str = "config.cfg", precision: int = PREC_FP64, overrides=None):
    """Reads a reference-style case directory: cfg plus input/particles.dat, input/objects.dat and
    input/body.dat (when present)."""
    with open(os.path.join(case_dir, cfg_name)) as fh:
        vals = config_values(fh.read(), overrides)
    scale = float(vals.get("scale", 1.0))
    ps = read_particles(os.path.join(case_dir, "input", "particles.dat"), scale)
    obs = read_objects(os.path.join(case_dir, "input", "objects.dat"), scale)
    body = os.path.join(case_dir, "input", "body.dat")
    if os.path.exists(body):
        obs.bodies = read_bodies(body, scale)
    return vals, params_from_values(vals, precision), ps, obs
