/*
 * lbdem.h — C ABI of the coupled D3Q19 free-surface lattice-Boltzmann / DEM timestep.
 *
 * Two libraries export exactly this interface:
 *   lbm-dem_amd/lib/liblbdem.so        — the MI355X (gfx950) implementation (HIP kernels)
 *   oracle/_build/liblbdem_refcpu.so   — the FP64 CPU oracle (test infrastructure only)
 *
 * The entry points replace the method surface that the reference's main loop calls
 * (reference src/main.cpp:38-67 goCycle, src/main.cpp:262-286 initialisation):
 *   lbdem_init            <- LB::latticeDefinition (LB.cpp:85-126),
 *                            DEM::discreteElementInit (DEM.cpp:199-357),
 *                            LB::latticeBolzmannInit (LB.cpp:340-398)
 *   lbdem_dem_step        <- DEM::discreteElementStep (DEM.cpp:359-401)
 *   lbdem_coupling_step   <- LB::latticeBoltzmannCouplingStep (LB.cpp:432-467)
 *   lbdem_lb_step         <- LB::latticeBolzmannStep (LB.cpp:400-418)
 *   lbdem_fs_step         <- LB::latticeBoltzmannFreeSurfaceStep (LB.cpp:422-430)
 *   lbdem_cycle           <- goCycle (main.cpp:38-67) repeated, minus IO::outputStep
 * Configuration values are the ones LB::latticeBoltzmannGet (LB.cpp:128-338),
 * DEM::discreteElementGet (DEM.cpp:17-197) and parseConfigFile (main.cpp:128-185) read
 * from the getpot file; particle/object arrays are the contents of input/particles.dat and
 * input/objects.dat after the reader's transformations (DEM.cpp:81-165).
 *
 * Conventions: plain pointers and sizes; every call returns 0 on success and a non-zero
 * LBDEM_E* code otherwise (lbdem_last_error gives the message). All arrays are host memory
 * owned by the caller and copied before the call returns. Axes follow the reference:
 * y is vertical. Cell index = x + y*X + z*X*Y (LB.cpp:2113-2137).
 */
#ifndef LBDEM_H
#define LBDEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBDEM_OK 0
#define LBDEM_E_ARG 1          /* invalid argument / unsupported configuration */
#define LBDEM_E_STATE 2        /* call out of order (e.g. step before init) */
#define LBDEM_E_DEVICE 3       /* HIP runtime error */
#define LBDEM_E_CAPACITY 4     /* a fixed capacity (neighbours, lists, scratch) overflowed */
#define LBDEM_E_TYPE 5         /* streaming met a neighbour type it cannot handle (LB.cpp:1524) */
#define LBDEM_E_NOMEM 6

#define LBDEM_PREC_FP64 0      /* populations stored and computed in FP64 */
#define LBDEM_PREC_FP32 1      /* populations stored as f - w_i in FP32, FP32 arithmetic */
#define LBDEM_PREC_FP16 2      /* populations stored as scaled (f - w_i) in FP16, FP32 arithmetic */

/* Parsed configuration. Physical units exactly as in the getpot file. */
typedef struct lbdem_params {
    /* main.cpp:131-175 */
    double demInitialRepeat;
    double maxTime;
    int demSolve, lbSolve, freeSurface, forceField;
    /* LB.cpp:137-336 */
    double unitLength, unitTime, unitDensity;
    double lbSize[3];            /* lbSizeX/Y/Z (physical) */
    double initVelocity[3];      /* physical */
    double lbF[3];               /* physical acceleration */
    int boundary[6];             /* boundary0..5: 4 periodic, 5-8 wall; every wall is of boundary0's
                                    kind (DEM.cpp:430-446): 7 static, 8 moving with zero velocity,
                                    5 static slip, 6 moving slip (slipCoefficient) */
    double slipCoefficient;      /* LB.cpp:313 */
    double fluidHeight;          /* physical */
    double initVisc;             /* physical dynamic viscosity (NEWTONIAN, FRICTIONAL, VOELLMY; the
                                    reference keeps its default (1.0 - 0.5) / 3.0 for BINGHAM) */
    /* rheology and Smagorinsky turbulence (LB.cpp:197-245, node.cpp:113-168) */
    int rheologyModel;           /* LBDEM_RHEO_*: 0 NEWTONIAN, 1 BINGHAM, 2 FRICTIONAL, 3 VOELLMY */
    double plasticVisc;          /* physical dynamic viscosity (BINGHAM) */
    double yieldStress;          /* physical stress (BINGHAM) */
    double frictionCoefFluid;    /* FRICTIONAL, VOELLMY */
    double voellmyConstant;      /* physical acceleration (VOELLMY) */
    int turbulenceSolve;         /* Smagorinsky term added to the apparent viscosity */
    double turbConst;            /* lattice units (the reference does not scale it) */
    /* DEM.cpp:19-75 */
    double density, objectDensity, bodyDensity, linearStiff, frictionCoefPart;
    double lubNormal, lubTangential, lubRotation, lubCutoff, kinVisc;
    double electroForce, debyeLength;
    int objWaitTime;
    /* body-spring impactor (DEM.cpp:63-73; bodyObjectInteraction DEM.cpp:798-869) */
    double springDamp, epsilon, k13;
    int nonLinType;              /* 0 spring, 1 + dashpot, 2 + cubic, 3 + repulsive LJ, 4 + WCA */
    /* implementation options (no reference counterpart) */
    int precision;               /* LBDEM_PREC_* (the CPU oracle always uses FP64) */
    int device;                  /* HIP device ordinal */
} lbdem_params;

/* rheologyModel values (node.h:17) */
enum { LBDEM_RHEO_NEWTONIAN = 0, LBDEM_RHEO_BINGHAM = 1, LBDEM_RHEO_FRICTIONAL = 2, LBDEM_RHEO_VOELLMY = 3 };

/* Per-particle record layout used by lbdem_get_particles (doubles). */
enum {
    LBDEM_P_X0 = 0, LBDEM_P_X1 = 3, LBDEM_P_W1 = 6, LBDEM_P_FHYDRO = 9, LBDEM_P_MHYDRO = 12,
    LBDEM_P_FCONTACT = 15, LBDEM_P_MCONTACT = 18, LBDEM_P_FLUB = 21, LBDEM_P_MLUB = 24,
    LBDEM_P_FREP = 27, LBDEM_P_FTOTAL = 30, LBDEM_P_MTOTAL = 33, LBDEM_P_FGRAV = 36,
    LBDEM_P_FLUIDVOL = 39, LBDEM_P_XIW = 40, LBDEM_P_XIO = 43, LBDEM_P_R = 46, LBDEM_P_M = 47,
    LBDEM_P_REC = 48
};
/* Per-object (impactor) record layout used by lbdem_get_objects (doubles). */
enum {
    LBDEM_O_X0 = 0, LBDEM_O_X1 = 3, LBDEM_O_W1 = 6, LBDEM_O_FHYDRO = 9, LBDEM_O_MHYDRO = 12,
    LBDEM_O_FCONTACT = 15, LBDEM_O_MCONTACT = 18, LBDEM_O_FLUB = 21, LBDEM_O_MLUB = 24,
    LBDEM_O_FTOTAL = 27, LBDEM_O_MTOTAL = 30, LBDEM_O_FGRAV = 33, LBDEM_O_R = 36, LBDEM_O_M = 37,
    LBDEM_O_REC = 38
};
/* Per-body record layout used by lbdem_get_bodies (doubles; body, elmt.h, DEM.cpp:168-196, 255-266). */
enum {
    LBDEM_B_X0 = 0, LBDEM_B_X1 = 3, LBDEM_B_FSPRING = 6, LBDEM_B_FTOTAL = 9, LBDEM_B_FGRAV = 12,
    LBDEM_B_L = 15, LBDEM_B_L0 = 16, LBDEM_B_K = 17, LBDEM_B_R = 18, LBDEM_B_M = 19,
    LBDEM_B_REC = 20
};
/* Per-wall record layout used by lbdem_get_walls (doubles). */
enum {
    LBDEM_W_FHYDRO = 0, LBDEM_W_FCONTACT = 3, LBDEM_W_FLUB = 6, LBDEM_W_FTOTAL = 9,
    LBDEM_W_REC = 12
};

typedef struct lbdem_info {
    int size[3];                 /* lbSize in cells (LB.cpp:170-172) */
    int64_t totNodes;
    int nParticles, nObjects, nWalls;
    int nActive, nFluid, nInterface, nParticleNodes, nObjectNodes;
    unsigned int lbTime;         /* LB::time */
    unsigned int demTimeStep;    /* DEM::demTimeStep */
    double demTime, realTime;
    int neighborListBuilds;      /* DEM::neighborList calls so far (incl. init) */
    int newNeighborList;         /* DEM::newNeighborList */
    double totalMassInit;        /* LB::totalMass (LB.cpp:388-395) */
    double lastExtraMass;        /* streaming extraMass of the last LB step (LB.cpp:1533) */
    double lastSurplus;          /* massSurplus of the last free-surface step (LB.cpp:1644) */
    int precision;
    int nBodies;
} lbdem_info;

typedef struct lbdem_ctx lbdem_ctx;

int lbdem_create(const lbdem_params* params, lbdem_ctx** out);
int lbdem_destroy(lbdem_ctx* ctx);
const char* lbdem_last_error(const lbdem_ctx* ctx);

/* particles.dat rows (DEM.cpp:88-134) after scaling: size, r*scale, x0*scale-0.05 (3n). */
int lbdem_set_particles(lbdem_ctx* ctx, int n, const int* size, const double* r, const double* x0);
/* objects.dat rows (DEM.cpp:138-165) after scaling: r, x0 (3n), x1 (3n). */
int lbdem_set_objects(lbdem_ctx* ctx, int n, const double* r, const double* x0, const double* x1);
/* body.dat rows (DEM.cpp:168-196) after scaling: r, x0*scale (3n), x1 (3n), spring stiffness k,
 * rest length L0; at most LBDEM_MAX_BODIES. Each body pulls every impactor along y (DEM.cpp:798-869)
 * from the impactor wait on, and moves under its gravity and spring forces (DEM.cpp:2728-2740). */
#define LBDEM_MAX_BODIES 8
int lbdem_set_bodies(lbdem_ctx* ctx, int n, const double* r, const double* x0, const double* x1, const double* k,
                     const double* L0);

/* Fluid restart from a fluidRecycle file (LB::latticeBolzmannInit with lbRestart, LB.cpp:352-367;
 * restartInterface LB.cpp:994-1009; restartVariables LB.cpp:1098-1138), between lbdem_set_objects
 * and lbdem_init; the host parses the file (IO.cpp:973-997 writes it). size: the lattice size the
 * file states (must equal the box's); types: one type code per node (index x + X (y + Y z)),
 * taken at every node whose initial type is fluid (codes 0, 2 and 3 are accepted there); records:
 * LBDEM_RESTART_REC values per record, n - 1, u (3), mass - 1, visc, f (19), consumed in node
 * index order by the active nodes of the restarted box (extra records are ignored, as the
 * reference ignores them). Each record's visc becomes the node's viscosity (node.cpp:40-50); a
 * NEWTONIAN run without turbulence keeps per-node viscosities only when some record's visc
 * differs from initVisc (in lattice units). */
#define LBDEM_RESTART_REC 25
int lbdem_set_restart_fluid(lbdem_ctx* ctx, const int size[3], const uint8_t* types, int nRecords,
                            const double* records);

int lbdem_init(lbdem_ctx* ctx);
/* Velocity and viscosity of every active node, as IO reads node->u and node->visc on an output
 * step (IO.cpp:973-997): the force-shifted velocity of the node's last collision (node.cpp:95-98,
 * LB.cpp:1310-1330), or the one the free surface gave a new interface node (LB.cpp:1705-1725).
 * The LB step after lbdem_store_velocity stores it for every active node (other steps keep it for
 * interface nodes only); lbdem_get_velocity returns u, 3 N values (x, y, z per node), and visc,
 * N values (0 where not active; either may be NULL), while no other LB step has run since (after
 * lbdem_init, the initial velocities), LBDEM_E_STATE otherwise. Not for contexts in a batch. */
int lbdem_store_velocity(lbdem_ctx* ctx);
int lbdem_get_velocity(lbdem_ctx* ctx, double* u, double* visc);
int lbdem_dem_step(lbdem_ctx* ctx);
int lbdem_coupling_step(lbdem_ctx* ctx);
int lbdem_lb_step(lbdem_ctx* ctx);
int lbdem_fs_step(lbdem_ctx* ctx);
/* nsteps repetitions of goCycle's solver part (main.cpp:38-64), including the
 * demInitialRepeat gating and the realTime/lb.time counters. */
int lbdem_cycle(lbdem_ctx* ctx, int nsteps);
int lbdem_synchronize(lbdem_ctx* ctx);

int lbdem_get_info(lbdem_ctx* ctx, lbdem_info* info);
/* type: node type code (node.h:122-135); pflag: inside-particle flag; solid: solidIndex. */
int lbdem_get_types(lbdem_ctx* ctx, uint8_t* type, uint8_t* pflag, int32_t* solid);
/* f: 19*N post-streaming populations, SoA (f[q*N+i]), 0 where the node does not exist;
 * mass, n: N values (0 where the node does not exist). Any pointer may be NULL. */
int lbdem_get_fields(lbdem_ctx* ctx, double* f, double* mass, double* n);
/* lbdem_get_fields without any allocation (no counterpart in the reference; for callers that
 * cannot allocate: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and values. The conversion goes through a 16 MiB device array that
 * each GPU holds from the library's load until the process ends, one field or population plane at
 * a time, so boxes of more than 2^21 (2,097,152) nodes are rejected with LBDEM_E_CAPACITY. Calls
 * on one GPU share that array: calls that overlap on one GPU race. */
int lbdem_get_fields_noalloc(lbdem_ctx* ctx, double* f, double* mass, double* n);
int lbdem_get_particles(lbdem_ctx* ctx, double* rec /* nParticles*LBDEM_P_REC */);
int lbdem_get_objects(lbdem_ctx* ctx, double* rec /* nObjects*LBDEM_O_REC */);
int lbdem_get_walls(lbdem_ctx* ctx, double* rec /* nWalls*LBDEM_W_REC */);
int lbdem_get_bodies(lbdem_ctx* ctx, double* rec /* nBodies*LBDEM_B_REC */);
/* Stress diagnostics (the DEM members totalStress, hydroStress, lubStress, contactStress,
 * electroStress and the particles' STotal and disp, DEM.cpp:655-771, 2705-2725, with the particle
 * stresslet SHydro of LB::computeHydroForces, LB.cpp:1951-1959). The reference computes them on
 * every step; here they are computed only when enabled, which lbdem_set_stress does before
 * lbdem_init (LBDEM_E_STATE after it). Once enabled, every LB step also stores the velocity of
 * every active node and runs one more pass over the particle nodes, and every DEM step
 * accumulates the pair stresses. Not available in batches. */
int lbdem_set_stress(lbdem_ctx* ctx, int enable);
/* The values of the last DEM step (as the reference's IO reads them after goCycle): global = 45
 * doubles, the tensors totalStress, hydroStress, lubStress, contactStress, electroStress in that
 * order, each m00 m01 m02 m10 ... m22, already divided by the volume (DEM.cpp:766-771);
 * stotal = nParticles*9 (STotal of each particle); disp = nParticles*3 (x_i - x0, DEM.cpp:2705).
 * Any pointer may be NULL. LBDEM_E_STATE unless enabled. */
int lbdem_get_stress(lbdem_ctx* ctx, double* global, double* stotal, double* disp);
/* Per-particle record of lbdem_get_particle_stress (doubles): the tensors of the particle VTK file
 * (STotal, SElas and SVisc with SVisc_h, DEM.cpp:745-761, 1091-1125, 1702-1773, 2103-2138), the
 * overlap FElas / linearStiff, disp, and the contact and normal-lubrication counts (DEM.cpp:1668,
 * 2362). */
enum {
    LBDEM_S_STOTAL = 0, LBDEM_S_SELAS = 9, LBDEM_S_SVISC = 18, LBDEM_S_SVISC_H = 27, LBDEM_S_OVERLAP = 36,
    LBDEM_S_DISP = 39, LBDEM_S_COORD = 42, LBDEM_S_LUB_COORD = 43, LBDEM_S_REC = 44
};
/* The last DEM step's per-particle stress records, nParticles*LBDEM_S_REC (stress diagnostics on). */
int lbdem_get_particle_stress(lbdem_ctx* ctx, double* rec);
/* Verlet list of every particle: count[n], list[n*cap + k] (DEM.cpp:473-652). The MI355X library
 * lists partners in ascending order, the oracle in the reference's traversal order; a pair that
 * the reference lists more than once (fewer than 3 link cells along z) appears as often. */
int lbdem_get_neighbors(lbdem_ctx* ctx, int cap, int* count, int* list);
/* Total fluid mass: sum of mass over active nodes not inside a particle (LB.cpp:388-395). */
int lbdem_total_mass(lbdem_ctx* ctx, double* mass);

/* Kernel timing (no reference counterpart; used by benchmarks): when enabled, HIP events on the
 * library's stream bracket every fused collide-stream launch; lbdem_profile_read returns the
 * summed duration in milliseconds and the number of launches since the last enable. */
int lbdem_profile_enable(lbdem_ctx* ctx, int enable);
int lbdem_profile_read(lbdem_ctx* ctx, double* fused_ms, int* fused_launches);

/* Device bytes of one context (no reference counterpart). lbdem_estimate_bytes computes, from
 * the parameters and the particle and impactor radii alone (no device call), the bytes
 * lbdem_init will allocate; lbdem_device_bytes reports what an initialised context holds. */
int lbdem_estimate_bytes(const lbdem_params* params, int nParticles, const double* r, int nObjects,
                         const double* objR, size_t* bytes);
int lbdem_device_bytes(lbdem_ctx* ctx, size_t* bytes);

/* Replay of captured step pairs in lbdem_cycle (no reference counterpart): on by default
 * (LBDEM_GRAPHS=0 in the environment turns the default off); 0 launches every step's kernels
 * one by one. Results are the same either way. */
int lbdem_use_graphs(lbdem_ctx* ctx, int enable);

/* Tile tuner of the fused collide-stream kernel (no reference counterpart). Tiles are 64 x 4
 * threads; each thread walks tile[2] nodes along z. Unless fixed, the first LB steps time the
 * candidates 8, 4, 2, 1 (one launch per AA parity each) on the run's own state and keep the
 * fastest; results do not depend on the tile beyond FP64 summation order of block partials.
 * lbdem_set_tile: 0 tunes (default), 1/2/4/8 fixes the z-walk length. */
typedef struct lbdem_tuning {
    int tile[3];
    int tuned;
    int nCandidates;
    int candidateZ[4];
    double candidateMs[4];       /* fused kernel time, both parities, per candidate */
    int pairCaptured;            /* 1 when lbdem_cycle's next step pair replays a captured graph,
                                    i.e. the context's start-up (tuning, capture) is over */
} lbdem_tuning;
int lbdem_set_tile(lbdem_ctx* ctx, int tileZ);
int lbdem_get_tuning(lbdem_ctx* ctx, lbdem_tuning* tuning);

/* ---------------------------------------------------------------------------------------------
 * Batch (no reference counterpart): several initialised contexts (same precision and device)
 * stepped together, every kernel of a step launched once for all of them (the launch's second
 * grid dimension selects the box; each box keeps its own memory, parameters and counters). A step
 * pair of the batch is captured into a HIP graph per phase and replayed. Members must have run
 * past their start-up steps (the impactor wait, the tile tuner, the first LB step) and be at the
 * same step parity; the batch runs every member with one z-walk length of the fused kernel (at
 * creation the largest box's; lbdem_batch_set_tile changes or tunes it). Results are those of
 * each context run alone with that walk. While a context belongs to a batch it is advanced only
 * through the batch; its downloads stay valid (they wait for the batch's work).
 * ------------------------------------------------------------------------------------------- */
typedef struct lbdem_batch lbdem_batch;
int lbdem_batch_create(lbdem_ctx* const* ctxs, int n, lbdem_batch** out);
int lbdem_batch_cycle(lbdem_batch* batch, int nsteps);  /* nsteps even */
int lbdem_batch_synchronize(lbdem_batch* batch);
const char* lbdem_batch_last_error(const lbdem_batch* batch);
int lbdem_batch_destroy(lbdem_batch* batch);         /* the contexts remain, usable alone again */
/* z-walk length for every member: 1, 2, 4 or 8; 0: the batch's next step pairs run one candidate
 * each (8, 4, 2), timed with events on the batch stream, and the fastest is kept (the tile
 * autotuner of a batch) */
int lbdem_batch_set_tile(lbdem_batch* batch, int tileZ);
/* the batch's walk; *tuned = 0 while candidates are still being timed */
int lbdem_batch_get_tile(const lbdem_batch* batch, int* tileZ, int* tuned);
/* Kernel timing of a batch (no reference counterpart; used by benchmarks): while enabled, each step
 * pair's batch launches run one by one with a HIP event after each (no graph replay; the host waits
 * for every pair). lbdem_batch_profile_read reports kernel i in order of first launch: its name
 * (valid until the next enable), summed milliseconds and launches since the last enable, and the
 * number of kernels in *count; an i past the end sets only *count and *kernel = NULL. */
int lbdem_batch_profile_enable(lbdem_batch* batch, int enable);
int lbdem_batch_profile_read(const lbdem_batch* batch, int i, const char** kernel, double* ms, int* launches,
                             int* count);

/* ---------------------------------------------------------------------------------------------
 * Run pool (no reference counterpart; batched and served runs of many independent boxes on one
 * GPU). A host thread admits submitted runs in arrival order while their estimated device bytes
 * fit the HBM budget and fewer than maxConcurrent runs are active, gives every run its own context
 * (memory carved from the pool's one allocation) and stream, advances the runs in chunks of
 * stepsPerChunk coupled steps (goCycle, as lbdem_cycle) -- alone during their start-up steps,
 * then together in batches -- and releases a run's device memory when its last step is done.
 * Runs never communicate: the result of a run does not depend on what else runs beside it.
 * ------------------------------------------------------------------------------------------- */
typedef struct lbdem_pool lbdem_pool;

typedef struct lbdem_run_desc {
    lbdem_params params;
    int nParticles;
    const int* size;             /* as lbdem_set_particles */
    const double* r;
    const double* x0;
    int nObjects;                /* as lbdem_set_objects */
    const double* objR;
    const double* objX0;
    const double* objX1;
    int nsteps;                  /* coupled steps to run */
} lbdem_run_desc;

#define LBDEM_RUN_QUEUED 0
#define LBDEM_RUN_ACTIVE 1
#define LBDEM_RUN_DONE 2
#define LBDEM_RUN_FAILED 3
#define LBDEM_RUN_CANCELLED 4

typedef struct lbdem_run_result {
    int state;                   /* LBDEM_RUN_* */
    int error;                   /* LBDEM_E* of a failed run */
    int stepsDone;
    int nObjects;
    double tSubmit, tStart, tDone;   /* seconds on the pool clock (lbdem_pool_clock) */
    double objects[2 * LBDEM_O_REC]; /* impactor records after the last step */
    double totalMass;            /* lbdem_total_mass after the last step */
    int nActive;                 /* active nodes after the last step */
    int tileZ;                   /* z-walk length of the fused kernel the run finished with */
    int64_t nodeSteps;           /* sum over steps of the box's node count (totNodes * steps) */
    size_t deviceBytes;
    char message[128];           /* error message of a failed run */
} lbdem_run_result;

/* hbmBudgetBytes: one device allocation the runs are carved from (0: 85% of the free memory);
 * maxConcurrent: most runs active at once; stepsPerChunk: coupled steps per enqueue (0: 32) */
int lbdem_pool_create(size_t hbmBudgetBytes, int maxConcurrent, int stepsPerChunk, int device, lbdem_pool** out);
/* copies the run's inputs; returns its id (0, 1, 2, ... in submission order) */
int lbdem_pool_submit(lbdem_pool* pool, const lbdem_run_desc* run, int* runId);
int lbdem_pool_poll(lbdem_pool* pool, int runId, lbdem_run_result* result);
/* waits until the run is done or failed, or timeoutSeconds passed (LBDEM_E_STATE then) */
int lbdem_pool_wait(lbdem_pool* pool, int runId, double timeoutSeconds, lbdem_run_result* result);
/* waits until every submitted run is done or failed */
int lbdem_pool_drain(lbdem_pool* pool, double timeoutSeconds);
/* a queued run is dropped; an active run stops after its chunks in flight; both end as
 * LBDEM_RUN_CANCELLED and release their memory. Other runs are not affected. */
int lbdem_pool_cancel(lbdem_pool* pool, int runId);

typedef struct lbdem_pool_stats {
    int submitted, queued, active, done, failed, cancelled;
    int activePeak;              /* most runs active at once */
    int batchBuilds;             /* batches formed (membership changes) */
    int batchedPeak;             /* most runs stepped by one batch */
    size_t budget, bytesInUse, bytesPeak;
} lbdem_pool_stats;
/* 1 (default): runs past their start-up steps are stepped together in batches (one per storage
 * precision; see lbdem_batch_create), re-formed between chunks as runs join and leave; 0: every
 * run on its own stream. Set before the first submission. */
int lbdem_pool_set_batching(lbdem_pool* pool, int enable);
/* z-walk of the runs' batches: 0 (default): tuned on the batches themselves
 * (lbdem_batch_set_tile(batch, 0)) per storage precision, again when a batch is at least twice the
 * size of the one tuned on (up to 8 runs), the result reused by later batches, and start-up steps
 * alone use the 8-node walk (with batching off each run tunes its own); 1, 2, 4 or 8: fixed for
 * every step of every run, batched or not. Set before the first submission. */
int lbdem_pool_set_batch_tile(lbdem_pool* pool, int tileZ);
int lbdem_pool_get_stats(lbdem_pool* pool, lbdem_pool_stats* stats);
double lbdem_pool_clock(const lbdem_pool* pool);
const char* lbdem_pool_last_error(const lbdem_pool* pool);
/* cancels what is queued or active, waits for the device and frees everything */
int lbdem_pool_destroy(lbdem_pool* pool);

#ifdef __cplusplus
}
#endif
#endif /* LBDEM_H */
