// C ABI of the MI355X library (include/lbdem.h): configuration, initial node types (the
// reference's initialisation order, LB.cpp:340-398 and DEM.cpp:199-357), device allocation,
// the goCycle step sequence (main.cpp:38-67) and downloads.
#include "lbdem_internal.hpp"
#include "launch.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <map>
#include <mutex>
#include <vector>

using namespace lbdem;

lbdem::Recorder*& lbdem::current_recorder() {
    thread_local Recorder* r = nullptr;
    return r;
}

namespace {

struct Units {  // conversion units (node.cpp:512-525)
    double Length = 1, Time = 1, Density = 1, Volume = 1, Speed = 1, Accel = 1, AngVel = 1, KinVisc = 1, DynVisc = 1,
           Force = 1, Torque = 1, Mass = 1, Stress = 1;
    void setComposite() {
        Volume = Length * Length * Length;
        Speed = Length / Time;
        Accel = Length / Time / Time;
        AngVel = 1.0 / Time;
        KinVisc = Length * Length / Time;
        DynVisc = Density * Length * Length / Time;
        Force = Density * Length * Length * Length * Length / Time / Time;
        Torque = Density * Length * Length * Length * Length * Length / Time / Time;
        Mass = Density * Length * Length * Length;
        Stress = Density * Length * Length / Time / Time;
    }
};

struct HostVec { double x = 0, y = 0, z = 0; };

}  // namespace

struct lbdem_ctx {
    lbdem_params prm{};
    std::string err;
    bool initialised = false;
    Units unit;
    int lbSize[3] = {1, 1, 1};
    double physSize[3] = {1, 1, 1};
    double bLoc[6][3];
    double lbF[3] = {0, 0, 0}, initVel[3] = {0, 0, 0};
    double initDynVisc = 0, rho0 = 1.0, fluidHeight = 0;
    // rheology in lattice units (LB.cpp:320-323); viscUpdate: the collision recomputes each node's
    // viscosity (non-Newtonian model or turbulence); viscPlane: nodes keep their own viscosity
    // (viscUpdate, or restart records with viscosities other than initDynVisc)
    double plasticVisc = 0, yieldStress = 0, voellmyCoef = 0;
    bool viscUpdate = false, viscPlane = false;
    bool dynWalls = false;  // wall nodes are reported as moving walls (type 8)
    bool slipWalls = false; // every wall slips (type 5 or 6): wall links get the slip pass
    bool slipDyn = false;   // moving slip walls (6): zero velocity, but in the free-surface mass exchange
    std::vector<int32_t> slipNodes;  // nodes with a diagonal link into a wall (init)
    bool freeSurface = false, forceField = false, lbFReset = false;
    // inputs
    std::vector<int> psize;
    std::vector<double> pr, px0, orad, ox0, ox1;
    std::vector<double> brad, bx0, bx1, bk, bL0;  // bodies (lbdem_set_bodies)
    // device
    hipStream_t stream = nullptr;
    // lbdem_cycle defers a step's free-surface pass to the start of the next step and runs it on
    // fsStream beside that step's DEM step (the two touch disjoint state); the coupling step
    // waits for both
    hipStream_t fsStream = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr, evHydro = nullptr;
    bool fsDeferred = false;
    bool streamError = false;  // an event record or wait of the overlap failed
    Box box{};
    Consts k{};
    DemParams dp{};
    DemState ds{};
    DevScalars* sc = nullptr;
    double* dTmp = nullptr;
    size_t dTmpBytes = 0;
    bool tmpPlanned = false;   // dTmp is the planned scratch buffer (freed with the rest)
    std::vector<void*> allocs;
    // device memory: separate hipMalloc per buffer, or carved from a region the caller owns
    char* region = nullptr;
    size_t regionSize = 0, regionUsed = 0;
    size_t devBytes = 0;
    // steady-state step pairs replayed from a captured HIP graph (see lbdem_cycle)
    // one captured pair per phase: AA parity x which history buffer is current
    bool useGraphs = true, graphFailed = false;
    hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t graphExec[4] = {nullptr, nullptr, nullptr, nullptr};
    double* xiFirst = nullptr;  // the buffer that is current at initialisation
    lbdem_batch* batch = nullptr;  // set while the context is advanced through a batch
    // tile tuner: the first LB steps run the fused kernel once per AA parity with each candidate
    // z-walk length, timed with events; the fastest candidate is kept (lbdem_set_tile fixes one)
    static constexpr int NCAND = 4;
    static constexpr int CAND[NCAND] = {8, 4, 2, 1};
    bool tuned = false;
    int tuneLaunches = 0;
    hipEvent_t tuneEv[2 * NCAND][2] = {};
    double tuneMs[NCAND] = {0, 0, 0, 0};
    // kernel timing
    bool profiling = false;
    std::vector<hipEvent_t> evPool;
    size_t evUsed = 0;
    // host counters (goCycle, main.cpp:40-43; DEM.cpp:365-366)
    unsigned lbTime = 0, demTimeStep = 0, currentTimeStep = 0;
    double demTime = 0, realTime = 0, deltat = 1.0, totalMass = 0;

    // fluid restart (lbdem_set_restart_fluid): the file's types and records, used by lbdem_init
    bool restart = false;
    std::vector<uint8_t> rTypes;
    std::vector<double> rRec;
    int rCount = 0;
    // node velocities (lbdem_store_velocity / lbdem_get_velocity): 0 not held, 1 the next LB step
    // stores them for every active node, 2 held (stored by the last LB step or by lbdem_init)
    int velState = 0;
    bool stress = false;  // stress diagnostics (lbdem_set_stress)

    const char* errFixed = nullptr;  // a message with static storage (set without allocating)
    int fail(int code, const std::string& m) { err = m; errFixed = nullptr; return code; }
    int fail_fixed(int code, const char* m) { errFixed = m; return code; }
    // 256-byte granules; from the region when one is set, else one hipMalloc per buffer
    void* raw_alloc(size_t bytes) {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        void* p = nullptr;
        if (region) {
            if (regionUsed + bytes > regionSize) return nullptr;
            p = region + regionUsed;
            regionUsed += bytes;
        } else {
            if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
            allocs.push_back(p);
        }
        devBytes += bytes;
        return p;
    }
    ~lbdem_ctx() {
        for (int q = 0; q < 4; ++q) {
            if (graphExec[q]) (void)hipGraphExecDestroy(graphExec[q]);
            if (graph[q]) (void)hipGraphDestroy(graph[q]);
        }
        for (hipEvent_t e : evPool) (void)hipEventDestroy(e);
        for (auto& pr : tuneEv)
            for (hipEvent_t e : pr)
                if (e) (void)hipEventDestroy(e);
        if (fsStream) (void)hipStreamSynchronize(fsStream);
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : allocs) (void)hipFree(p);
        if (dTmp && !tmpPlanned) (void)hipFree(dTmp);
        if (evFork) (void)hipEventDestroy(evFork);
        if (evJoin) (void)hipEventDestroy(evJoin);
        if (evHydro) (void)hipEventDestroy(evHydro);
        if (fsStream) (void)hipStreamDestroy(fsStream);
        if (stream) (void)hipStreamDestroy(stream);
    }
    // download scratch: the planned buffer, replaced by a hipMalloc of its own when a download
    // needs more (field downloads only)
    double* tmp(size_t bytes) {
        if (bytes > dTmpBytes) {
            if (dTmp && !tmpPlanned) (void)hipFree(dTmp);
            dTmp = nullptr;
            tmpPlanned = false;
            if (hipMalloc(&dTmp, bytes) != hipSuccess) { dTmpBytes = 0; return nullptr; }
            dTmpBytes = bytes;
        }
        return dTmp;
    }
};

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return c->fail(LBDEM_E_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

static int read_params(lbdem_ctx* c) {
    const lbdem_params& P = c->prm;
    Units& u = c->unit;
    u.Length = P.unitLength; u.Time = P.unitTime; u.Density = P.unitDensity;
    if (!(u.Length > 0) || !(u.Time > 0) || !(u.Density > 0)) return c->fail(LBDEM_E_ARG, "unit lengths/times must be > 0");
    for (int a = 0; a < 3; ++a) {
        if (!(P.lbSize[a] > 0)) return c->fail(LBDEM_E_ARG, "lbSize must be > 0");
        c->lbSize[a] = int(std::floor(P.lbSize[a] / u.Length + 0.5));  // LB.cpp:170-172
        if (c->lbSize[a] < 3) return c->fail(LBDEM_E_ARG, "lattice must have at least 3 nodes per axis");
        c->physSize[a] = double(unsigned(c->lbSize[a])) * u.Length;
    }
    const double L = u.Length;
    const double loc[6][3] = {{0.5 * L, 0, 0}, {double(unsigned(c->lbSize[0]) - 1.5) * L, 0, 0},
                              {0, 0.5 * L, 0}, {0, double(unsigned(c->lbSize[1]) - 1.5) * L, 0},
                              {0, 0, 0.5 * L}, {0, 0, double(unsigned(c->lbSize[2]) - 1.5) * L}};
    std::memcpy(c->bLoc, loc, sizeof(loc));
    if (!(P.initVisc > 0)) return c->fail(LBDEM_E_ARG, "initVisc must be > 0");
    u.setComposite();
    for (int a = 0; a < 3; ++a) {
        c->initVel[a] = P.initVelocity[a] / u.Speed;
        c->lbF[a] = P.lbF[a] / u.Accel;
    }
    for (int i = 0; i < 6; ++i)
        if (P.boundary[i] != 4 && (P.boundary[i] < 5 || P.boundary[i] > 8))
            return c->fail(LBDEM_E_ARG, "boundary codes must be 4 (periodic) or 5-8 (walls)");
    // every wall takes boundary0's kind (DEM.cpp:430-446 test externalBoundary[0] for all walls):
    // slip walls are not supported; a moving (8) wall keeps the zero velocity its nodes are given
    // (LB.cpp:1162-1169), so it reflects as a static wall and only its type code differs
    c->dynWalls = P.boundary[0] == 8;
    c->slipWalls = P.boundary[0] == 5 || P.boundary[0] == 6;
    c->slipDyn = P.boundary[0] == 6;
    if (c->slipWalls && !(P.slipCoefficient >= 0.0)) return c->fail(LBDEM_E_ARG, "slipCoefficient must be >= 0");
    c->fluidHeight = P.fluidHeight;
    c->initDynVisc = P.initVisc / u.DynVisc;
    if (P.rheologyModel < LBDEM_RHEO_NEWTONIAN || P.rheologyModel > LBDEM_RHEO_VOELLMY)
        return c->fail(LBDEM_E_ARG, "unknown rheologyModel");
    c->plasticVisc = P.plasticVisc / u.DynVisc;
    c->yieldStress = P.yieldStress / u.Stress;
    c->voellmyCoef = P.voellmyConstant / u.Accel;
    c->viscUpdate = P.rheologyModel != LBDEM_RHEO_NEWTONIAN || P.turbulenceSolve != 0;
    c->viscPlane = c->viscUpdate;
    c->freeSurface = P.freeSurface != 0;
    c->forceField = P.forceField != 0;
    if (!P.lbSolve) return c->fail(LBDEM_E_ARG, "lbSolve=0 is not supported");
    if (!(P.density > 0) || !(P.objectDensity > 0)) return c->fail(LBDEM_E_ARG, "density/objectDensity must be > 0");
    if (P.precision < 0 || P.precision > 2) return c->fail(LBDEM_E_ARG, "unknown precision");
    return 0;
}

extern "C" int lbdem_create(const lbdem_params* params, lbdem_ctx** out) {
    if (!params || !out) return LBDEM_E_ARG;
    lbdem_ctx* c = new lbdem_ctx();
    *out = c;
    c->prm = *params;
    return read_params(c);
}

extern "C" int lbdem_destroy(lbdem_ctx* ctx) {
    if (ctx) {
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    }
    return 0;
}

extern "C" const char* lbdem_last_error(const lbdem_ctx* ctx) {
    return ctx ? (ctx->errFixed ? ctx->errFixed : ctx->err.c_str()) : "null context";
}

extern "C" int lbdem_set_particles(lbdem_ctx* c, int n, const int* size, const double* r, const double* x0) {
    if (!c || n < 0 || (n > 0 && (!r || !x0))) return LBDEM_E_ARG;
    if (c->initialised) return c->fail(LBDEM_E_STATE, "particles must be set before lbdem_init");
    c->psize.assign(n, 1);
    if (size) for (int i = 0; i < n; ++i) c->psize[i] = size[i];
    c->pr.assign(r, r + n);
    c->px0.assign(x0, x0 + 3 * size_t(n));
    return 0;
}

extern "C" int lbdem_set_objects(lbdem_ctx* c, int n, const double* r, const double* x0, const double* x1) {
    if (!c || n < 0 || (n > 0 && (!r || !x0 || !x1))) return LBDEM_E_ARG;
    if (n > MAX_OBJECTS) return c->fail(LBDEM_E_ARG, "too many objects");
    if (n > 2) return c->fail(LBDEM_E_ARG, "at most 2 objects are supported by the lattice kernels");
    if (c->initialised) return c->fail(LBDEM_E_STATE, "objects must be set before lbdem_init");
    c->orad.assign(r, r + n);
    c->ox0.assign(x0, x0 + 3 * size_t(n));
    c->ox1.assign(x1, x1 + 3 * size_t(n));
    return 0;
}

extern "C" int lbdem_set_bodies(lbdem_ctx* c, int n, const double* r, const double* x0, const double* x1,
                                const double* k, const double* L0) {
    if (!c || n < 0 || (n > 0 && (!r || !x0 || !x1 || !k || !L0))) return LBDEM_E_ARG;
    if (n > LBDEM_MAX_BODIES) return c->fail(LBDEM_E_ARG, "too many bodies");
    if (c->initialised) return c->fail(LBDEM_E_STATE, "bodies must be set before lbdem_init");
    // nonLinType outside 0-4 leaves the reference's spring force undefined (DEM.cpp:801-860)
    if (n > 0 && (c->prm.nonLinType < 0 || c->prm.nonLinType > 4))
        return c->fail(LBDEM_E_ARG, "nonLinType must be 0-4 with bodies");
    c->brad.assign(r, r + n);
    c->bx0.assign(x0, x0 + 3 * size_t(n));
    c->bx1.assign(x1, x1 + 3 * size_t(n));
    c->bk.assign(k, k + n);
    c->bL0.assign(L0, L0 + n);
    return 0;
}

// LB::latticeBolzmannInit's restart branch (LB.cpp:352-367): the file's size must be the
// lattice's (the reference asserts it); the types and records are kept for lbdem_init
extern "C" int lbdem_set_restart_fluid(lbdem_ctx* c, const int size[3], const uint8_t* types, int nRecords,
                                       const double* records) {
    if (!c || !size || !types || nRecords < 0 || (nRecords > 0 && !records)) return LBDEM_E_ARG;
    if (c->initialised) return c->fail(LBDEM_E_STATE, "the restart must be set before lbdem_init");
    for (int a = 0; a < 3; ++a)
        if (size[a] != c->lbSize[a]) return c->fail(LBDEM_E_ARG, "the restart file's lattice size differs from the box's");
    const size_t n = size_t(c->lbSize[0]) * size_t(c->lbSize[1]) * size_t(c->lbSize[2]);
    c->rTypes.assign(types, types + n);
    c->rRec.assign(records, records + size_t(nRecords) * LBDEM_RESTART_REC);
    c->rCount = nRecords;
    c->restart = true;
    return 0;
}

// initial node types on the host: lattice boundaries (LB.cpp:538-564), DEM walls
// (LB.cpp:713-748), objects (LB.cpp:751-795), gas above the fluid height (LB.cpp:974-992) and
// interface nodes next to gas (LB.cpp:1058-1068)
// initial node types (initializeLatticeBoundaries, initializeTypes, initializeInterface;
// LB.cpp:340-398) and the impactor node list (initializeObjectBoundaries, LB.cpp:751-795). With a
// restart the file's type replaces every fluid node's (restartInterface, LB.cpp:994-1009) instead
// of the free-surface initialisation: fluid, gas, interface, or an impactor node (setObjectWall,
// node.cpp:405), which keeps solid index 0 and, not on the impactor list, is never uncovered (the
// file holds the impactor where it was when written). Returns false for any other type there.
static bool host_types(lbdem_ctx* c, const Geom& g, std::vector<uint8_t>& t, std::vector<int32_t>& s,
                       const std::vector<std::array<double, 6>>& walls, std::vector<int32_t>& objList) {
    const int X = g.X, Y = g.Y, Z = g.Z;
    t.assign(g.N, T_FLUID);
    s.assign(g.N, 0);
    const int* B = c->prm.boundary;
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                const int64_t i = cidx(g, x, y, z);
                int v = -1;
                if (x == 0) v = B[0]; else if (x == X - 1) v = B[1];
                else if (y == 0) v = B[2]; else if (y == Y - 1) v = B[3];
                else if (z == 0) v = B[4]; else if (z == Z - 1) v = B[5];
                if (v >= 0) t[i] = uint8_t(v);
            }
    for (size_t w = 0; w < walls.size(); ++w) {
        const double px = walls[w][0] / c->unit.Length, py = walls[w][1] / c->unit.Length, pz = walls[w][2] / c->unit.Length;
        const double nx = walls[w][3], ny = walls[w][4], nz = walls[w][5];
        for (int z = 0; z < Z; ++z)
            for (int y = 0; y < Y; ++y)
                for (int x = 0; x < X; ++x) {
                    const double d = nx * (x - px) + ny * (y - py) + nz * (z - pz);
                    if (d < 0.0) { const int64_t i = cidx(g, x, y, z); s[i] = int32_t(w); t[i] = T_WALL; }
                }
    }
    for (size_t o = 0; o < c->orad.size(); ++o) {
        const double cx = c->ox0[3 * o] / c->unit.Length, cy = c->ox0[3 * o + 1] / c->unit.Length, cz = c->ox0[3 * o + 2] / c->unit.Length;
        const double r = c->orad[o] / c->unit.Length;
        for (int z = 0; z < Z; ++z)
            for (int y = 0; y < Y; ++y)
                for (int x = 0; x < X; ++x) {
                    const double dx = x - cx, dy = y - cy, dz = z - cz;
                    if (dx * dx + dy * dy + dz * dz < r * r) { const int64_t i = cidx(g, x, y, z); s[i] = int32_t(o); t[i] = T_OBJ; }
                }
    }
    objList.clear();
    for (int64_t i = 0; i < g.N; ++i) if (t[i] == T_OBJ) objList.push_back(int32_t(i));
    if (c->restart) {
        for (int64_t i = 0; i < g.N; ++i) {
            if (t[i] != T_FLUID) continue;
            const uint8_t v = c->rTypes[size_t(i)];
            if (v != T_FLUID && v != T_GAS && v != T_IF && v != T_OBJ) return false;
            t[i] = v;
        }
    } else {
        if (c->freeSurface) {
            const double h0 = c->fluidHeight / c->unit.Length;
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y)
                    for (int x = 0; x < X; ++x) {
                        const int64_t i = cidx(g, x, y, z);
                        if (unsigned(y) > h0 && t[i] == T_FLUID) t[i] = T_GAS;
                    }
        }
        for (int z = 0; z < Z; ++z)
            for (int y = 0; y < Y; ++y)
                for (int x = 0; x < X; ++x) {
                    const int64_t i = cidx(g, x, y, z);
                    if (t[i] != T_FLUID) continue;
                    for (int q = 1; q < Q; ++q)
                        if (t[nbr_wrapped(g, x, y, z, q)] == T_GAS) { t[i] = T_IF; break; }
                }
    }
    // wall nodes carry the wall that receives their link forces (LB.cpp:1416-1420)
    const int nw = int(walls.size());
    for (int64_t i = 0; i < g.N; ++i)
        if (t[i] == T_WALL) {
            const int w = (s[i] >= 0 && s[i] < nw) ? s[i] : 0;
            t[i] = uint8_t(T_WALL | (w << T_WSHIFT));
        }
    return true;
}

// ---------------------------------------------------------------------------------------------
// device buffers of a context: one list walked both by lbdem_init (allocating) and by
// lbdem_estimate_bytes (summing), so the estimate is the bytes init takes
// ---------------------------------------------------------------------------------------------
constexpr size_t TMP_PLANNED = 8192;  // scratch of lbdem_total_mass / lbdem_get_info

struct Plan {
    Geom g;
    int np = 0, no = 0, nw = 0, ncell = 1;
    int nl[3] = {0, 0, 0};
    size_t selt = 8, celt = 8;
    int ifCap = 0, objCap = 0;
    bool visc = false;  // per-node viscosity plane
    int nSlip = 0;      // slip-pass nodes
    bool stress = false;  // stress diagnostic buffers
};

static int wall_count(const lbdem_ctx* c) {
    int n = 0;
    for (int i = 0; i < 6; ++i) n += c->prm.boundary[i] >= 5 && c->prm.boundary[i] <= 8;
    return n;
}

// DEM linked cells (DEM.cpp:505-554): cell edge 2*rmax+2 over the DEM box; the direct sum (all
// zero here) when x or y has fewer than 3 cells. z is not tested (DEM.cpp:515): with 1 or 2
// z-cells the wrapped 27-cell walk visits a z-cell more than once and lists such pairs once per
// visit, and with 0 z-cells no list is built at all; both are kept.
static void dem_cells(const lbdem_ctx* c, double rmax, int nl[3]) {
    const double demSize[3] = {c->physSize[0], c->freeSurface ? c->fluidHeight : c->physSize[1], c->physSize[2]};
    const double cut = 2.0 * rmax + 2.0;
    for (int a = 0; a < 3; ++a) nl[a] = int(demSize[a] / cut);
    if (std::min(nl[0], nl[1]) < 3) nl[0] = nl[1] = nl[2] = 0;
}

static int make_plan(lbdem_ctx* c, Plan& p) {
    Geom& g = p.g;
    g.X = c->lbSize[0]; g.Y = c->lbSize[1]; g.Z = c->lbSize[2];
    g.N = int64_t(g.X) * g.Y * g.Z;
    if (g.N >= (int64_t(1) << 31) / Q) return c->fail(LBDEM_E_ARG, "box too large for 32-bit population indexing");
    for (int i = 0; i < 6; ++i) g.per[i] = c->prm.boundary[i] == 4;
    p.np = int(c->pr.size());
    p.no = int(c->orad.size());
    p.nw = wall_count(c);
    double rmax = 0.0;
    for (double r : c->pr) rmax = r > rmax ? r : rmax;
    dem_cells(c, rmax, p.nl);
    p.ncell = std::max(1, p.nl[0] * p.nl[1] * p.nl[2]);
    const int prec = c->prm.precision;
    p.selt = prec == LBDEM_PREC_FP64 ? 8 : (prec == LBDEM_PREC_FP32 ? 4 : 2);
    p.celt = prec == LBDEM_PREC_FP64 ? 8 : 4;
    p.ifCap = int(std::max<int64_t>(65536, g.N / 4));
    int64_t objCells = 0;
    for (int o = 0; o < p.no; ++o) {
        const double r = c->orad[o] / c->unit.Length + 4.0;
        objCells += int64_t(std::ceil(8.0 * r * r * r));
    }
    p.objCap = int(std::min<int64_t>(std::max<int64_t>(objCells, 1024), g.N));
    p.visc = c->viscPlane;
    p.stress = c->stress;
    if (c->slipWalls) {  // at most the two node layers next to each wall face
        const int64_t area[3] = {int64_t(g.Y) * g.Z, int64_t(g.X) * g.Z, int64_t(g.X) * g.Y};
        int64_t cnt = 0;
        for (int i = 0; i < 6; ++i)
            if (c->prm.boundary[i] >= 5 && c->prm.boundary[i] <= 8) cnt += 2 * area[i / 2];
        p.nSlip = int(std::min<int64_t>(std::max<int64_t>(cnt, 1), g.N));
    }
    return 0;
}

// calls alloc(void*&, bytes) for every device buffer of the context; returns false when one
// of them came back null
template <class T> struct EltSize { static constexpr size_t value = sizeof(T); };
template <> struct EltSize<void> { static constexpr size_t value = 1; };  // byte buffers

template <class F>
static bool plan_buffers(const Plan& p, Box& b, DemState& d, DevScalars*& sc, double*& tmp, F&& alloc) {
    bool ok = true;
    auto A = [&](auto*& ptr, size_t count) {
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(ptr)>>;
        void* v = nullptr;
        alloc(v, std::max<size_t>(count, 1) * EltSize<T>::value);
        ptr = static_cast<T*>(v);
        ok = ok && v != nullptr;
    };
    const size_t N = size_t(p.g.N);
    const size_t np1 = size_t(std::max(p.np, 1)), no1 = size_t(std::max(p.no, 1));
    b.ifCap = p.ifCap;
    b.objCap = p.objCap;
    b.changedCap = b.ifCap + b.objCap;
    b.blockPartCap = std::max(fused_grid_blocks(p.g, 1, LBDEM_PREC_FP32), fused_grid_blocks(p.g, 1, LBDEM_PREC_FP16)) * PART_STRIDE_;  // the largest grid (z-walk 1)
    A(b.f, size_t(Q) * N * p.selt);
    A(b.type, N);
    A(b.solid, N);
    A(b.mass, N * p.celt);
    A(b.aux, 7 * N * p.celt);
    A(b.ifList, size_t(b.ifCap));
    A(b.ifList2, size_t(b.ifCap) + 16);
    A(b.ifNewMass, size_t(b.ifCap) * p.celt);
    A(b.filled, size_t(b.ifCap));
    A(b.emptied, size_t(b.ifCap));
    A(b.newIf, size_t(b.ifCap));
    A(b.donor, N);
    A(b.flagNew, N);
    A(b.objList, size_t(b.objCap));
    A(b.objCand, size_t(b.objCap));
    A(b.coverCount, COVER_BLOCKS + UNCOVER_BLOCKS_MAX);
    A(b.changed, size_t(b.changedCap));
    A(b.blockPart, size_t(b.blockPartCap));
    A(b.blockFlag, size_t(std::max(fused_grid_blocks(p.g, 1, LBDEM_PREC_FP32), fused_grid_blocks(p.g, 1, LBDEM_PREC_FP16))));
    A(b.bcm, N);
    if (p.visc) A(b.visc, N * p.celt);
    else b.visc = nullptr;
    if (p.nSlip > 0) A(b.slipList, size_t(p.nSlip));
    else b.slipList = nullptr;
    b.nSlip = p.nSlip;
    A(b.hacc, np1 * 8);
    A(sc, 1);
    // DEM
    const size_t n3 = 3 * np1;
    double** vecs[] = {&d.x0, &d.x1, &d.w1, &d.FHydro, &d.MHydro, &d.FContact, &d.MContact, &d.FLub, &d.MLub, &d.FRep,
                       &d.FTotal, &d.MTotal, &d.FGrav, &d.xiw, &d.xio};
    for (double** v : vecs) A(*v, n3);
    double** pvecs[] = {&d.r, &d.m, &d.I, &d.im, &d.iI, &d.fluidVol};
    for (double** v : pvecs) A(*v, np1);
    const size_t slots = np1 * NBR_CAP;
    A(d.nbrCount, np1); A(d.nbrCountOld, np1);
    A(d.nbrList, slots); A(d.nbrListOld, slots); A(d.nbrSlot, slots);
    A(d.xiCur, 3 * slots); A(d.xiNext, 3 * slots);
    A(d.cellCount, size_t(p.ncell)); A(d.cellStart, size_t(p.ncell)); A(d.cellFill, size_t(p.ncell));
    A(d.cellItems, np1); A(d.cellKey, np1);
    A(d.objContrib, np1 * 12 * no1);
    A(d.wallContrib, np1 * 6 * size_t(std::max(p.nw, 1)));
    A(d.objSum, 12 * MAX_OBJECTS);
    A(d.redPart, 2 * 64 * 12 * MAX_OBJECTS);
    A(d.wallSum, 6 * MAX_WALLS);
    A(d.blockMax, VEL_SLOTS);
    double** ovecs[] = {&d.ox0, &d.ox1, &d.ow1, &d.oFHydro, &d.oMHydro, &d.oFContact, &d.oMContact, &d.oFLub,
                        &d.oMLub, &d.oFTotal, &d.oMTotal, &d.oFGrav};
    for (double** v : ovecs) A(*v, 3 * no1);
    double** opvecs[] = {&d.or_, &d.om, &d.oI, &d.oim, &d.oiI};
    for (double** v : opvecs) A(*v, no1);
    double** wvecs[] = {&d.wFHydro, &d.wFContact, &d.wFLub, &d.wFTotal};
    for (double** v : wvecs) A(*v, 3 * size_t(MAX_WALLS));
    A(d.body, size_t(LBDEM_MAX_BODIES) * LBDEM_B_REC);
    if (p.stress) {
        A(d.stressPart, np1 * SP_REC); A(d.shydro, np1 * 9); A(d.xi0, n3); A(d.disp, n3);
    } else {
        d.stressPart = d.shydro = d.xi0 = d.disp = nullptr;
    }
    A(d.plb, np1);
    A(d.olb, no1);
    A(tmp, TMP_PLANNED / sizeof(double));
    return ok;
}

static size_t plan_bytes(const Plan& p) {
    Box b{};
    DemState d{};
    DevScalars* sc = nullptr;
    double* tmp = nullptr;
    size_t total = 0;
    plan_buffers(p, b, d, sc, tmp, [&](void*& v, size_t bytes) {
        total += (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        v = &total;  // any non-null marker: nothing is dereferenced
    });
    return total;
}

extern "C" int lbdem_estimate_bytes(const lbdem_params* params, int nParticles, const double* r, int nObjects,
                                    const double* objR, size_t* bytes) {
    if (!params || !bytes || nParticles < 0 || nObjects < 0 || (nParticles > 0 && !r) || (nObjects > 0 && !objR))
        return LBDEM_E_ARG;
    lbdem_ctx c;
    c.prm = *params;
    if (int e = read_params(&c)) return e;
    c.pr.assign(r, r + nParticles);
    c.orad.assign(objR, objR + nObjects);
    Plan p;
    if (int e = make_plan(&c, p)) return e;
    *bytes = plan_bytes(p);
    return 0;
}

extern "C" int lbdem_device_bytes(lbdem_ctx* c, size_t* bytes) {
    if (!c || !bytes) return LBDEM_E_ARG;
    *bytes = c->devBytes;
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Download scratch of lbdem_get_fields_noalloc: a file-scope device array, so that entry point
// calls no allocation. Each GPU holds its own copy from the library's load until the process
// ends. In host code the array's name is only a stand-in for it: its device address is looked up
// once per device (with that device current, in lbdem_init) and kept in scratchPtr; every host-side
// use (kernel arguments, copies) takes that pointer.
// ---------------------------------------------------------------------------------------------
constexpr size_t NOALLOC_DOUBLES = size_t(1) << 21;  // 16 MiB: one field of a box of up to 2^21 nodes
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_DEVICES = 64;
static std::mutex scratchMu;
static double* scratchPtr[NOALLOC_DEVICES];

// the calling thread's current device must be ``dev``; returns a hipError_t
static int noalloc_scratch_setup(int dev) {
    if (dev < 0 || dev >= NOALLOC_DEVICES) return int(hipErrorInvalidDevice);
    std::lock_guard<std::mutex> lk(scratchMu);
    if (scratchPtr[dev]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return int(e != hipSuccess ? e : hipErrorInvalidSymbol);
    scratchPtr[dev] = static_cast<double*>(ptr);
    return 0;
}
static double* noalloc_scratch(int dev) {
    if (dev < 0 || dev >= NOALLOC_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(scratchMu);
    return scratchPtr[dev];
}

extern "C" int lbdem_init(lbdem_ctx* c) {
    if (!c) return LBDEM_E_ARG;
    if (c->initialised) return c->fail(LBDEM_E_STATE, "already initialised");
    HIPCHK(hipSetDevice(c->prm.device));
    if (const int e = noalloc_scratch_setup(c->prm.device); e != 0)
        return c->fail(LBDEM_E_DEVICE, std::string("hipGetSymbolAddress of the download scratch failed: ") +
                                           hipGetErrorString(hipError_t(e)));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->fsStream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->evHydro, hipEventDisableTiming));
    if (const char* g = std::getenv("LBDEM_GRAPHS")) c->useGraphs = std::atoi(g) != 0;
    if (!c->tuned)
        for (auto& pr : c->tuneEv)
            for (hipEvent_t& e : pr) HIPCHK(hipEventCreate(&e));
    const Units& u = c->unit;
    const int np = int(c->pr.size()), no = int(c->orad.size());
    if (np > (1 << 30)) return c->fail(LBDEM_E_ARG, "too many particles");

    // ---------------- DEM set-up (DEM.cpp:199-357, elmt.cpp:17-53) ----------------
    DemParams& P = c->dp;
    std::memset(&P, 0, sizeof(P));
    P.demSize[0] = c->physSize[0];
    P.demSize[1] = c->freeSurface ? c->fluidHeight : c->physSize[1];
    P.demSize[2] = c->physSize[2];
    double demF[3];
    for (int a = 0; a < 3; ++a) demF[a] = c->lbF[a] * u.Accel;
    std::vector<double> m(np), I(np), im(np), iI(np), fg(3 * size_t(np));
    double meanR = 0.0, rmax = 0.0;
    for (int i = 0; i < np; ++i) {
        const double r = c->pr[i];
        const double single = 4.0 / 3.0 * c->prm.density * M_PI * r * r * r;
        m[i] = c->psize[i] * single;
        I[i] = c->psize[i] * 2.0 / 5.0 * single * r * r;
        im[i] = 1.0 / m[i];
        iI[i] = 1.0 / I[i];
        for (int a = 0; a < 3; ++a) fg[3 * i + a] = demF[a] * m[i];
        meanR += r;
        rmax = r > rmax ? r : rmax;
    }
    if (np > 0) meanR /= double(np);
    std::vector<double> om(no), oI(no), oim(no), oiI(no), ofg(3 * size_t(no));
    for (int o = 0; o < no; ++o) {
        const double r = c->orad[o];
        om[o] = 4.0 / 3.0 * c->prm.objectDensity * M_PI * r * r * r;
        for (int a = 0; a < 3; ++a) ofg[3 * o + a] = demF[a] * om[o];
        oI[o] = 2.0 / 5.0 * om[o] * r * r;
        oim[o] = 1.0 / om[o];
        oiI[o] = 1.0 / oI[o];
    }
    // bodies (DEM.cpp:257-266): mass, gravity, L = L0; the records as lbdem_get_bodies returns them
    const int nbody = int(c->brad.size());
    std::vector<double> brec(size_t(std::max(nbody, 1)) * LBDEM_B_REC, 0.0);
    for (int b = 0; b < nbody; ++b) {
        double* R = &brec[size_t(b) * LBDEM_B_REC];
        const double r = c->brad[b];
        const double m = c->prm.bodyDensity * (4.0 / 3.0) * M_PI * (r * r * r);
        for (int a = 0; a < 3; ++a) {
            R[LBDEM_B_X0 + a] = c->bx0[3 * b + a];
            R[LBDEM_B_X1 + a] = c->bx1[3 * b + a];
            R[LBDEM_B_FGRAV + a] = demF[a] * m;
        }
        R[LBDEM_B_L] = c->bL0[b]; R[LBDEM_B_L0] = c->bL0[b]; R[LBDEM_B_K] = c->bk[b];
        R[LBDEM_B_R] = r; R[LBDEM_B_M] = m;
    }
    std::vector<std::array<double, 6>> walls;
    const double dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int i = 0; i < 6; ++i)
        if (c->prm.boundary[i] >= 5 && c->prm.boundary[i] <= 8)
            walls.push_back({c->bLoc[i][0], c->bLoc[i][1], c->bLoc[i][2], dirs[i][0], dirs[i][1], dirs[i][2]});
    c->deltat = u.Time / double(1);
    const double avgMass = 4.0 / 3.0 * c->prm.density * M_PI * meanR * meanR * meanR;
    P.deltat = c->deltat;
    P.linearStiff = c->prm.linearStiff;
    P.dampCoeff = std::sqrt(avgMass * c->prm.linearStiff);
    P.frictionCoefPart = c->prm.frictionCoefPart;
    P.kinVisc = c->prm.kinVisc;
    P.lubNormal = c->prm.lubNormal; P.lubTangential = c->prm.lubTangential;
    P.lubRotation = c->prm.lubRotation; P.lubCutoff = c->prm.lubCutoff;
    P.nBody = nbody;
    P.nonLinType = c->prm.nonLinType;
    P.springDamp = c->prm.springDamp; P.epsilon = c->prm.epsilon; P.k13 = c->prm.k13;
    P.electroForce = c->prm.electroForce; P.debyeLength = c->prm.debyeLength;
    P.rmax = rmax;
    P.objWaitTime = c->prm.objWaitTime;
    for (int a = 0; a < 3; ++a) P.periodic[a] = c->prm.boundary[2 * a] == 4 && c->prm.boundary[2 * a + 1] == 4;
    Plan plan;
    if (int e = make_plan(c, plan)) return e;
    P.nlx = plan.nl[0]; P.nly = plan.nl[1]; P.nlz = plan.nl[2];
    P.nPart = np; P.nObj = no; P.nWall = int(walls.size());
    for (size_t w = 0; w < walls.size(); ++w)
        for (int a = 0; a < 3; ++a) { P.wallP[w][a] = walls[w][a]; P.wallN[w][a] = walls[w][3 + a]; }
    P.unitLength = u.Length; P.unitSpeed = u.Speed; P.unitAngVel = u.AngVel;
    P.stress = c->stress ? 1 : 0;

    // ---------------- lattice ----------------
    const Geom g = plan.g;
    std::vector<uint8_t> t;
    std::vector<int32_t> s, objList;
    if (c->restart && (c->rTypes.size() != size_t(g.N)))
        return c->fail(LBDEM_E_ARG, "the restart file's lattice size differs from the box's");
    if (!host_types(c, g, t, s, walls, objList))
        return c->fail(LBDEM_E_ARG, "the restart file puts a type other than fluid, gas, interface or impactor at a fluid node");
    // slip walls: every node that is neither wall nor periodic layer with a diagonal link into a
    // wall (walls do not change; the slip pass checks activity every step)
    c->slipNodes.clear();
    if (c->slipWalls) {
        for (int z = 0; z < g.Z; ++z)
            for (int y = 0; y < g.Y; ++y)
                for (int x = 0; x < g.X; ++x) {
                    const int64_t i = cidx(g, x, y, z);
                    const int base = t[i] & T_BASE;
                    if (base == T_WALL || base == T_PER) continue;
                    for (int j = 7; j < Q; ++j)
                        if ((t[nbr_wrapped(g, x, y, z, j)] & T_BASE) == T_WALL) { c->slipNodes.push_back(int32_t(i)); break; }
                }
        if (int64_t(c->slipNodes.size()) > plan.nSlip) return c->fail(LBDEM_E_STATE, "slip node list exceeds its plan");
    }
    // restartVariables (LB.cpp:1098-1138): one record per active node, in node index order
    std::vector<int32_t> rNode;
    if (c->restart) {
        for (int64_t i = 0; i < g.N; ++i)
            if (t[i] == T_FLUID || t[i] == T_IF) rNode.push_back(int32_t(i));
        if (int64_t(rNode.size()) > c->rCount)
            return c->fail(LBDEM_E_ARG, "the restart file holds " + std::to_string(c->rCount) + " records for " +
                                            std::to_string(rNode.size()) + " active nodes");
        // a record with another viscosity than the run's: nodes keep their own (node.cpp:40-50)
        for (size_t r = 0; r < rNode.size() && !c->viscPlane; ++r)
            if (c->rRec[r * LBDEM_RESTART_REC + 5] != c->initDynVisc) c->viscPlane = true;
        plan.visc = c->viscPlane;
    }

    Box& b = c->box;
    b.g = g;
    b.prec = c->prm.precision;
    const size_t selt = plan.selt, celt = plan.celt;
    b.parity = 0;
    DemState& d = c->ds;
    double* tmpBuf = nullptr;
    if (!plan_buffers(plan, b, d, c->sc, tmpBuf, [&](void*& v, size_t bytes) { v = c->raw_alloc(bytes); }))
        return c->fail(LBDEM_E_NOMEM, "device allocation failed");
    c->dTmp = tmpBuf;
    c->dTmpBytes = TMP_PLANNED;
    c->tmpPlanned = true;
    hipStream_t st = c->stream;
    HIPCHK(hipMemsetAsync(b.f, 0, size_t(Q) * g.N * selt, st));
    HIPCHK(hipMemsetAsync(b.mass, 0, g.N * celt, st));
    // particle hydrodynamic sums: zero before the first LB step (k_hydro_finalize re-zeroes them)
    HIPCHK(hipMemsetAsync(b.hacc, 0, size_t(std::max(int(c->pr.size()), 1)) * 8 * sizeof(double), st));
    // largest particle |v|^2 of the step, for the neighbour-list trigger: velocities start at zero
    HIPCHK(hipMemsetAsync(c->ds.blockMax, 0, VEL_SLOTS * sizeof(unsigned long long), st));
    // DEM cell list: each rebuild clears it after use (k_nbr_slots), the first one finds it zero
    HIPCHK(hipMemsetAsync(c->ds.cellCount, 0, size_t(plan.ncell) * sizeof(int), st));
    HIPCHK(hipMemsetAsync(c->ds.cellFill, 0, size_t(plan.ncell) * sizeof(int), st));
    HIPCHK(hipMemsetAsync(b.aux, 0, 7 * g.N * celt, st));
    HIPCHK(hipMemsetAsync(b.donor, 0xff, g.N * 4, st));
    HIPCHK(hipMemsetAsync(b.flagNew, 0xff, g.N * 4, st));
    HIPCHK(hipMemcpyAsync(b.type, t.data(), g.N, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b.solid, s.data(), g.N * 4, hipMemcpyHostToDevice, st));
    if (int64_t(objList.size()) > b.objCap) return c->fail(LBDEM_E_CAPACITY, "object node list capacity");
    if (!objList.empty()) HIPCHK(hipMemcpyAsync(b.objList, objList.data(), objList.size() * 4, hipMemcpyHostToDevice, st));
    DevScalars hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.nObjList = int(objList.size());
    hs.kminP = -1;
    hs.newNeighborList = 1;
    hs.fullMark = 1;

    // DEM device state
    const size_t n3 = 3 * size_t(std::max(np, 1));
    const size_t slots = size_t(std::max(np, 1)) * NBR_CAP;
    const size_t o3 = 3 * size_t(std::max(no, 1));
    const size_t w3 = 3 * size_t(MAX_WALLS);
    double* vecs[] = {d.x0, d.x1, d.w1, d.FHydro, d.MHydro, d.FContact, d.MContact, d.FLub, d.MLub, d.FRep,
                      d.FTotal, d.MTotal, d.FGrav, d.xiw, d.xio};
    for (double* v : vecs) HIPCHK(hipMemsetAsync(v, 0, n3 * 8, st));
    HIPCHK(hipMemsetAsync(d.fluidVol, 0, std::max(np, 1) * 8, st));
    double* ovecs[] = {d.ox0, d.ox1, d.ow1, d.oFHydro, d.oMHydro, d.oFContact, d.oMContact, d.oFLub,
                       d.oMLub, d.oFTotal, d.oMTotal, d.oFGrav};
    for (double* v : ovecs) HIPCHK(hipMemsetAsync(v, 0, o3 * 8, st));
    double* wvecs[] = {d.wFHydro, d.wFContact, d.wFLub, d.wFTotal};
    for (double* v : wvecs) HIPCHK(hipMemsetAsync(v, 0, w3 * 8, st));
    HIPCHK(hipMemsetAsync(d.nbrCount, 0, std::max(np, 1) * 4, st));
    HIPCHK(hipMemsetAsync(d.nbrCountOld, 0, std::max(np, 1) * 4, st));
    HIPCHK(hipMemsetAsync(d.xiCur, 0, 3 * slots * 8, st));
    HIPCHK(hipMemsetAsync(d.xiNext, 0, 3 * slots * 8, st));
    if (np > 0) {
        HIPCHK(hipMemcpyAsync(d.x0, c->px0.data(), 3 * np * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.r, c->pr.data(), np * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.m, m.data(), np * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.I, I.data(), np * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.im, im.data(), np * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.iI, iI.data(), np * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.FGrav, fg.data(), 3 * np * 8, hipMemcpyHostToDevice, st));
    }
    if (no > 0) {
        HIPCHK(hipMemcpyAsync(d.ox0, c->ox0.data(), 3 * no * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.ox1, c->ox1.data(), 3 * no * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.or_, c->orad.data(), no * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.om, om.data(), no * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.oI, oI.data(), no * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.oim, oim.data(), no * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.oiI, oiI.data(), no * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d.oFGrav, ofg.data(), 3 * no * 8, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(d.body, brec.data(), brec.size() * 8, hipMemcpyHostToDevice, st));
    if (c->stress) {  // x_i = x0 at the start (DEM.cpp:288); SHydro zero until the first LB step
        HIPCHK(hipMemsetAsync(d.stressPart, 0, size_t(std::max(np, 1)) * SP_REC * 8, st));
        HIPCHK(hipMemsetAsync(d.shydro, 0, size_t(std::max(np, 1)) * 9 * 8, st));
        HIPCHK(hipMemsetAsync(d.disp, 0, n3 * 8, st));
        HIPCHK(hipMemsetAsync(d.xi0, 0, n3 * 8, st));
        if (np > 0) HIPCHK(hipMemcpyAsync(d.xi0, c->px0.data(), 3 * np * 8, hipMemcpyHostToDevice, st));
    }
    b.nSlip = int(c->slipNodes.size());
    if (b.nSlip > 0)
        HIPCHK(hipMemcpyAsync(b.slipList, c->slipNodes.data(), size_t(b.nSlip) * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->sc, &hs, sizeof(hs), hipMemcpyHostToDevice, st));

    // constants of the lattice kernels
    Consts& k = c->k;
    for (int a = 0; a < 3; ++a) { k.lbF[a] = c->lbF[a]; k.initVel[a] = c->initVel[a]; }
    k.visc = c->initDynVisc;
    k.rho0 = c->rho0;
    k.omega = 1.0 / (0.5 + 3.0 * k.visc);
    k.omegaf = 1.0 - 1.0 / (1.0 + 6.0 * k.visc);
    k.rheo = c->prm.rheologyModel;
    k.turb = c->prm.turbulenceSolve != 0 ? 1 : 0;
    k.viscUpdate = c->viscUpdate ? 1 : 0;
    k.plasticVisc = c->plasticVisc;
    k.yieldStress = c->yieldStress;
    k.frictionCoef = c->prm.frictionCoefFluid;
    k.voellmyCoef = c->voellmyCoef;
    k.turbConst = c->prm.turbConst;
    k.slip = c->slipDyn ? 2 : (c->slipWalls ? 1 : 0);
    k.slipCoef = c->prm.slipCoefficient;
    k.minVisc = (0.500001 - 0.5) / 3.0;  // lattice.h:26-27 (minTau, maxTau), node.cpp:116-117
    k.maxVisc = (1.8 - 0.5) / 3.0;
    k.freeSurface = c->freeSurface ? 1 : 0;
    k.nPart = np; k.nObj = no; k.nWall = int(walls.size());
    k.unitLength = u.Length; k.unitSpeed = u.Speed; k.unitAngVel = u.AngVel;
    k.unitForce = u.Force; k.unitTorque = u.Torque; k.unitVolume = u.Volume;

    c->xiFirst = d.xiCur;
    dem_refresh_views(P, d, st);
    dem_neighbor_list(P, d, c->sc, true, st);   // DEM.cpp:294
    if (c->restart) {
        const size_t nr = rNode.size();
        double* dRec = nullptr;
        int32_t* dNode = nullptr;
        if (nr > 0) {
            if (hipMalloc(&dRec, nr * LBDEM_RESTART_REC * sizeof(double)) != hipSuccess ||
                hipMalloc(&dNode, nr * sizeof(int32_t)) != hipSuccess) {
                if (dRec) (void)hipFree(dRec);
                return c->fail(LBDEM_E_NOMEM, "device allocation of the restart records failed");
            }
            HIPCHK(hipMemcpyAsync(dRec, c->rRec.data(), nr * LBDEM_RESTART_REC * sizeof(double), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(dNode, rNode.data(), nr * sizeof(int32_t), hipMemcpyHostToDevice, st));
            lb_restart_fields(b, dRec, dNode, int(nr), st);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        if (dRec) (void)hipFree(dRec);
        if (dNode) (void)hipFree(dNode);
        c->rTypes.clear(); c->rTypes.shrink_to_fit();
        c->rRec.clear(); c->rRec.shrink_to_fit();
    } else {
        lb_init_fields(b, k, st);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    c->velState = 2;  // the initial (or restarted) velocity of every active node is in aux
    // total mass (LB.cpp:388-395): no node is flagged yet, all active nodes count
    double total = 0.0;
    {
        const int nb = 1024;
        double* part = c->tmp(nb * sizeof(double));
        if (!part) return c->fail(LBDEM_E_NOMEM, "device allocation failed");
        lb_total_mass(b, part, st);
        std::vector<double> h(nb);
        HIPCHK(hipMemcpyAsync(h.data(), part, nb * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (double v : h) total += v;
    }
    c->totalMass = total;
    c->lbTime = 0; c->demTime = 0.0; c->realTime = 0.0; c->demTimeStep = 0;
    c->initialised = true;
    return 0;
}

static int check_errors(lbdem_ctx* c) {
    DevScalars hs;
    HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (hs.typeError) return c->fail(LBDEM_E_TYPE, "TYPE ERROR: streaming met an unsupported neighbour type");
    if (hs.listOverflow) return c->fail(LBDEM_E_CAPACITY, "an interface/object list overflowed its capacity");
    if (hs.nbrOverflow) return c->fail(LBDEM_E_CAPACITY, "neighbour list capacity (64) exceeded");
    return 0;
}

// --------------------------- steps ---------------------------
// hydroReady: the previous LB step's hydrodynamic sums are finalised on another stream; the
// integration (the first DEM launch that reads them) waits for it
static void enqueue_dem(lbdem_ctx* c, hipEvent_t hydroReady = nullptr) {  // DEM::discreteElementStep (DEM.cpp:359-401)
    c->demTimeStep++;
    c->demTime += c->deltat;
    dem_neighbor_list(c->dp, c->ds, c->sc, false, c->stream);
    if (!dem_step(c->dp, c->ds, c->sc, int(c->demTimeStep), c->stream, hydroReady)) c->streamError = true;
}
static void enqueue_coupling(lbdem_ctx* c) {  // LB::latticeBoltzmannCouplingStep (LB.cpp:432-467)
    // the lists of the rebuild resolve overlaps unless the reference builds none (no z-cell)
    const bool lists = c->dp.nlx == 0 || c->dp.nlz > 0;
    const bool cells = c->dp.nlx > 0 && c->dp.nlz > 0;  // the rebuild sorted the particles by link cell
    lb_flags_rebuild_compute(c->box, c->k, c->ds.plb, c->sc, c->ds.nbrCount, c->ds.nbrList, lists,
                             cells ? c->ds.cellItems : nullptr, c->stream);
    lb_object_update(c->box, c->k, c->ds.olb, c->sc, c->stream);
    lb_flags_rebuild_apply(c->box, c->sc, c->stream);
    lb_find_new_active(c->box, c->k, c->ds.plb, c->sc, c->stream);
}
// picks the fastest candidate from the timed launches (waits for the last of them)
static void finish_tuning(lbdem_ctx* c) {
    int best = 0;
    double bestMs = 1e300;
    bool ok = hipEventSynchronize(c->tuneEv[2 * lbdem_ctx::NCAND - 1][1]) == hipSuccess;
    for (int q = 0; ok && q < lbdem_ctx::NCAND; ++q) {
        double ms = 0.0;
        for (int p = 0; p < 2; ++p) {
            float t = 0.0f;
            ok = ok && hipEventElapsedTime(&t, c->tuneEv[2 * q + p][0], c->tuneEv[2 * q + p][1]) == hipSuccess;
            ms += t;
        }
        c->tuneMs[q] = ms;
        if (ms < bestMs) { bestMs = ms; best = q; }
    }
    if (!ok) (void)hipGetLastError();
    c->box.fz = lbdem_ctx::CAND[ok ? best : 0];
    c->tuned = true;
}

// the particle, impactor and wall sums of the fused kernel (and the counters of the free-surface
// pass that follows)
static void enqueue_hydro(lbdem_ctx* c, hipStream_t s) {
    lb_hydro_finalize(c->box, c->k, c->ds.FHydro, c->ds.MHydro, c->ds.fluidVol, c->sc, c->ds.oFHydro, c->ds.oMHydro,
                      c->ds.wFHydro, true, s);
}
// finalize = false: the hydrodynamic finalisation is left to the next step's fork (it goes with
// the deferred free surface)
static void enqueue_lb(lbdem_ctx* c, bool finalize = true) {  // LB::latticeBolzmannStep (LB.cpp:400-418)
    if (!c->forceField && !c->lbFReset) {  // LB.cpp:1303-1305
        c->k.lbF[0] = c->k.lbF[1] = c->k.lbF[2] = 0.0;
        c->lbFReset = true;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->profiling) {
        while (c->evPool.size() < c->evUsed + 2) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) break;
            c->evPool.push_back(e);
        }
        if (c->evPool.size() >= c->evUsed + 2) { e0 = c->evPool[c->evUsed]; e1 = c->evPool[c->evUsed + 1]; c->evUsed += 2; }
    }
    if (!c->tuned && !c->profiling) {
        if (c->tuneLaunches < 2 * lbdem_ctx::NCAND) {  // next candidate launch, timed
            c->box.fz = lbdem_ctx::CAND[c->tuneLaunches / 2];
            e0 = c->tuneEv[c->tuneLaunches][0];
            e1 = c->tuneEv[c->tuneLaunches][1];
            ++c->tuneLaunches;
        } else {
            finish_tuning(c);
        }
    }
    lb_mark_edges(c->box, c->sc, c->stream);
    const bool storeVel = c->velState == 1;  // lbdem_store_velocity asked for this step
    c->velState = storeVel ? 2 : 0;
    // the stress diagnostics read every active node's hydrodynamic force from aux
    lb_fused_step(c->box, c->k, c->ds.plb, c->ds.olb, c->sc, true, storeVel || c->stress, c->stream, e0, e1);
    if (c->stress)
        lb_hydro_stress(c->box, c->ds.plb, c->dp.nPart, c->unit.Length, c->unit.Force, c->ds.shydro, c->stream);
    if (finalize) enqueue_hydro(c, c->stream);
}
static void enqueue_fs(lbdem_ctx* c, hipStream_t s) { lb_free_surface(c->box, c->k, c->sc, s); }

#define NEED_INIT                                                           \
    if (!c) return LBDEM_E_ARG;                                             \
    if (!c->initialised) return c->fail(LBDEM_E_STATE, "call lbdem_init first")

extern "C" int lbdem_dem_step(lbdem_ctx* c) { NEED_INIT; if (c->batch) return c->fail(LBDEM_E_STATE, "in a batch"); enqueue_dem(c); HIPCHK(hipGetLastError()); return 0; }
extern "C" int lbdem_coupling_step(lbdem_ctx* c) { NEED_INIT; if (c->batch) return c->fail(LBDEM_E_STATE, "in a batch"); enqueue_coupling(c); HIPCHK(hipGetLastError()); return 0; }
extern "C" int lbdem_lb_step(lbdem_ctx* c) { NEED_INIT; if (c->batch) return c->fail(LBDEM_E_STATE, "in a batch"); enqueue_lb(c); HIPCHK(hipGetLastError()); return 0; }
extern "C" int lbdem_fs_step(lbdem_ctx* c) {
    NEED_INIT;
    if (c->batch) return c->fail(LBDEM_E_STATE, "in a batch");
    if (c->freeSurface) enqueue_fs(c, c->stream);
    HIPCHK(hipGetLastError());
    return 0;
}

// one goCycle (main.cpp:38-67) minus the output. With deferFs the step's free-surface pass is
// left for the next step (or flush_fs), which runs it on fsStream beside its DEM step: the free
// surface reads and writes the lattice (populations, types, mass, interface lists) and the DEM
// step the particles and the hydrodynamic sums the LB step left, so the order between them does
// not matter; the coupling step waits for both, as it follows both in goCycle.
static void enqueue_step(lbdem_ctx* c, bool deferFs = false) {
    c->realTime += c->unit.Time;
    ++c->currentTimeStep;
    ++c->lbTime;
    bool join = false;
    hipEvent_t hydroReady = nullptr;
    if (c->fsDeferred) {  // the previous step's hydrodynamic sums and free surface (LB.cpp:422-430)
        c->fsDeferred = false;
        if (c->prm.demSolve) {
            if (hipEventRecord(c->evFork, c->stream) != hipSuccess ||
                hipStreamWaitEvent(c->fsStream, c->evFork, 0) != hipSuccess)
                c->streamError = true;
            enqueue_hydro(c, c->fsStream);
            if (hipEventRecord(c->evHydro, c->fsStream) != hipSuccess) c->streamError = true;
            hydroReady = c->evHydro;
            enqueue_fs(c, c->fsStream);
            if (hipEventRecord(c->evJoin, c->fsStream) != hipSuccess) c->streamError = true;
            join = true;
        } else {
            enqueue_hydro(c, c->stream);
            enqueue_fs(c, c->stream);
        }
    }
    if (c->prm.demSolve) enqueue_dem(c, hydroReady);
    if (join && hipStreamWaitEvent(c->stream, c->evJoin, 0) != hipSuccess) c->streamError = true;
    if (c->demTime >= c->prm.demInitialRepeat) {
        if (c->prm.demSolve) enqueue_coupling(c);
        const bool defer = deferFs && c->freeSurface;
        enqueue_lb(c, !defer);
        if (c->freeSurface) {
            if (defer) c->fsDeferred = true;
            else enqueue_fs(c, c->stream);
        }
    } else if (c->demTime <= c->prm.demInitialRepeat) {
        if (c->prm.demSolve) enqueue_coupling(c);
    }
}
// the deferred hydrodynamic finalisation and free-surface pass, on the context's stream (every
// lbdem_cycle ends with them)
static void flush_fs(lbdem_ctx* c) {
    if (!c->fsDeferred) return;
    c->fsDeferred = false;
    enqueue_hydro(c, c->stream);
    enqueue_fs(c, c->stream);
}
// lbdem_cycle overlaps a step's free surface with the next DEM step
static bool overlap_fs(const lbdem_ctx* c) { return c->freeSurface && c->prm.demSolve; }

// A pair of steps has the same launches with the same arguments as every later pair once
//  * every step compares its DEM step number past objWaitTime (the only use of the number in
//    the kernels: DEM.cpp:716-730, 777-783, 2727),
//  * the LB path runs (demTime >= demInitialRepeat, which stays true),
//  * the one-off reset of the lattice force (LB.cpp:1303-1305) is done;
// the AA parity and the tangential-history buffers alternate with period two. Every
// data-dependent branch is a device flag, so the pair can be captured once and replayed.
static bool steady_pairs(const lbdem_ctx* c) {
    if (!c->useGraphs || c->graphFailed || c->profiling || !c->tuned) return false;
    if (c->velState == 1) return false;  // the next LB step stores every node's velocity
    if (c->prm.demSolve && c->demTimeStep < unsigned(std::max(c->prm.objWaitTime, 0)) + 1u) return false;
    if (!(c->demTime >= c->prm.demInitialRepeat)) return false;
    if (!c->forceField && !c->lbFReset) return false;
    return true;
}
static bool graph_ready(const lbdem_ctx* c) {
    if (!steady_pairs(c)) return false;
    if (c->fsDeferred != overlap_fs(c)) return false;  // a pair starts with the previous free surface
    return true;
}

struct HostStepState {
    unsigned lbTime, demTimeStep, currentTimeStep;
    double demTime, realTime;
    double* xiCur;
    double* xiNext;
    int parity;
    bool fsDeferred;
};
static HostStepState save_host(const lbdem_ctx* c) {
    return {c->lbTime, c->demTimeStep, c->currentTimeStep, c->demTime, c->realTime, c->ds.xiCur, c->ds.xiNext,
            c->box.parity, c->fsDeferred};
}
static void restore_host(lbdem_ctx* c, const HostStepState& h) {
    c->lbTime = h.lbTime; c->demTimeStep = h.demTimeStep; c->currentTimeStep = h.currentTimeStep;
    c->demTime = h.demTime; c->realTime = h.realTime;
    c->ds.xiCur = h.xiCur; c->ds.xiNext = h.xiNext; c->box.parity = h.parity;
    c->fsDeferred = h.fsDeferred;
}

// the launches of a step pair depend on the phase: AA parity and current history buffer
static int graph_phase(const lbdem_ctx* c) { return c->box.parity * 2 + (c->ds.xiCur == c->xiFirst ? 0 : 1); }

// captures two steps into c->graphExec[phase] and launches them: 1 on success; 0 when capture
// or instantiation failed (host state restored, graphs switched off for this context, nothing
// executed); -1 when the launch of the captured steps failed
static int capture_pair(lbdem_ctx* c) {
    const int ph = graph_phase(c);
    const HostStepState h = save_host(c);
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        c->graphFailed = true;
        return 0;
    }
    const bool defer = overlap_fs(c);
    enqueue_step(c, defer);
    enqueue_step(c, defer);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    hipGraphExec_t ge = nullptr;
    if (e != hipSuccess || !g || c->streamError || hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        restore_host(c, h);
        c->streamError = false;
        c->graphFailed = true;
        return 0;
    }
    c->graph[ph] = g;
    c->graphExec[ph] = ge;
    // the host counters now describe the two captured steps; run them
    return hipGraphLaunch(ge, c->stream) == hipSuccess ? 1 : -1;
}

// what two enqueue_step calls do to the host counters (AA parity and history buffers return to
// their values after a pair)
static void advance_pair(lbdem_ctx* c) {
    c->velState = 0;  // two LB steps without the velocity store
    for (int k = 0; k < 2; ++k) {
        c->realTime += c->unit.Time;
        ++c->currentTimeStep;
        ++c->lbTime;
        if (c->prm.demSolve) { ++c->demTimeStep; c->demTime += c->deltat; }
    }
}

extern "C" int lbdem_cycle(lbdem_ctx* c, int nsteps) {
    NEED_INIT;
    if (c->batch) return c->fail(LBDEM_E_STATE, "the context is advanced through its batch");
    int s = 0;
    while (s < nsteps) {
        if (nsteps - s >= 2 && graph_ready(c)) {
            if (hipGraphExec_t ge = c->graphExec[graph_phase(c)]) {
                HIPCHK(hipGraphLaunch(ge, c->stream));
                advance_pair(c);
                s += 2;
                continue;
            }
            const int r = capture_pair(c);
            if (r < 0) {
                flush_fs(c);
                return c->fail(LBDEM_E_DEVICE, "launch of the captured step pair failed");
            }
            if (r == 1) {
                s += 2;
                continue;
            }
        }
        enqueue_step(c, overlap_fs(c));
        ++s;
    }
    flush_fs(c);  // every call leaves whole steps behind
    if (c->streamError) {
        c->streamError = false;
        return c->fail(LBDEM_E_DEVICE, "ordering of the free-surface stream failed");
    }
    HIPCHK(hipGetLastError());
    return 0;
}

static void drop_graphs(lbdem_ctx* c) {
    for (int q = 0; q < 4; ++q) {
        if (c->graphExec[q]) (void)hipGraphExecDestroy(c->graphExec[q]);
        if (c->graph[q]) (void)hipGraphDestroy(c->graph[q]);
        c->graphExec[q] = nullptr;
        c->graph[q] = nullptr;
    }
}

extern "C" int lbdem_set_tile(lbdem_ctx* c, int tileZ) {
    if (!c) return LBDEM_E_ARG;
    if (tileZ != 0 && tileZ != 1 && tileZ != 2 && tileZ != 4 && tileZ != 8)
        return c->fail(LBDEM_E_ARG, "tile z-walk length must be 0 (tune), 1, 2, 4 or 8");
    if (c->stream) HIPCHK(hipStreamSynchronize(c->stream));
    drop_graphs(c);  // captured steps hold the previous tile
    if (tileZ == 0) {
        if (c->initialised && c->tuned)
            for (auto& pr : c->tuneEv)
                for (hipEvent_t& e : pr)
                    if (!e) HIPCHK(hipEventCreate(&e));
        c->tuned = false;
        c->tuneLaunches = 0;
    } else {
        c->box.fz = tileZ;
        c->tuned = true;
    }
    return 0;
}

extern "C" int lbdem_get_tuning(lbdem_ctx* c, lbdem_tuning* t) {
    if (!c || !t) return LBDEM_E_ARG;
    std::memset(t, 0, sizeof(*t));
    t->tile[0] = FUSED_TILE_X;
    t->tile[1] = FUSED_TILE_Y;
    t->tile[2] = c->box.fz;
    t->tuned = c->tuned ? 1 : 0;
    t->nCandidates = lbdem_ctx::NCAND;
    for (int q = 0; q < lbdem_ctx::NCAND; ++q) {
        t->candidateZ[q] = lbdem_ctx::CAND[q];
        t->candidateMs[q] = c->tuneMs[q];
    }
    // Between calls the deferred free-surface pass has been flushed, so an lbdem_cycle call takes
    // its first step alone and replays pairs from the phase one step later (AA parity and, with
    // particles, the history buffer flip every step).
    int next = graph_phase(c);
    if (!c->fsDeferred && overlap_fs(c)) next ^= 2 | ((c->prm.demSolve && !c->pr.empty()) ? 1 : 0);
    t->pairCaptured = steady_pairs(c) && c->graphExec[next] ? 1 : 0;
    return 0;
}

extern "C" int lbdem_use_graphs(lbdem_ctx* c, int enable) {
    if (!c) return LBDEM_E_ARG;
    c->useGraphs = enable != 0;
    return 0;
}

// hooks of the run pool (pool.cpp)
int lbdem::ctx_use_region(lbdem_ctx* c, void* base, size_t bytes) {
    if (!c || !base) return LBDEM_E_ARG;
    if (c->initialised || c->devBytes) return c->fail(LBDEM_E_STATE, "the memory region must be set before lbdem_init");
    c->region = static_cast<char*>(base);
    c->regionSize = bytes;
    c->regionUsed = 0;
    return 0;
}
hipStream_t lbdem::ctx_stream(const lbdem_ctx* c) { return c ? c->stream : nullptr; }

extern "C" int lbdem_synchronize(lbdem_ctx* c) {
    NEED_INIT;
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_errors(c);
}

// --------------------------- downloads ---------------------------
extern "C" int lbdem_get_info(lbdem_ctx* c, lbdem_info* in) {
    NEED_INIT;
    std::memset(in, 0, sizeof(*in));
    for (int a = 0; a < 3; ++a) in->size[a] = c->lbSize[a];
    in->totNodes = c->box.g.N;
    in->nParticles = c->dp.nPart; in->nObjects = c->dp.nObj; in->nWalls = c->dp.nWall;
    in->nBodies = c->dp.nBody;
    int* cnt = (int*)c->tmp(64);
    if (!cnt) return c->fail(LBDEM_E_NOMEM, "device allocation failed");
    HIPCHK(hipMemsetAsync(cnt, 0, 5 * sizeof(int), c->stream));
    lb_count_types(c->box, cnt, c->stream);
    int h[5];
    DevScalars hs;
    HIPCHK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    in->nFluid = h[0]; in->nInterface = h[1]; in->nObjectNodes = h[2]; in->nParticleNodes = h[3];
    in->nActive = h[0] + h[1];
    in->lbTime = c->lbTime; in->demTimeStep = c->demTimeStep; in->demTime = c->demTime; in->realTime = c->realTime;
    in->neighborListBuilds = hs.builds;
    in->newNeighborList = hs.newNeighborList;
    in->totalMassInit = c->totalMass;
    in->lastExtraMass = hs.lastExtraMass;
    in->lastSurplus = hs.lastSurplus;
    in->precision = c->box.prec;
    return check_errors(c);
}

extern "C" int lbdem_get_types(lbdem_ctx* c, uint8_t* type, uint8_t* pflag, int32_t* solid) {
    NEED_INIT;
    const int64_t N = c->box.g.N;
    std::vector<uint8_t> t(N);
    HIPCHK(hipMemcpyAsync(t.data(), c->box.type, N, hipMemcpyDeviceToHost, c->stream));
    if (solid) HIPCHK(hipMemcpyAsync(solid, c->box.solid, N * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint8_t wallCode = c->dynWalls ? 8 : (c->slipDyn ? 6 : (c->slipWalls ? 5 : T_WALL));
    for (int64_t i = 0; i < N; ++i) {
        if (type) type[i] = (t[i] & T_BASE) == T_WALL ? wallCode : uint8_t(t[i] & T_BASE);
        if (pflag) pflag[i] = (t[i] & T_PMASK) ? 1 : 0;
    }
    return 0;
}

extern "C" int lbdem_store_velocity(lbdem_ctx* c) {
    NEED_INIT;
    if (c->batch) return c->fail(LBDEM_E_STATE, "the context is advanced through its batch");
    c->velState = 1;
    return 0;
}

extern "C" int lbdem_get_velocity(lbdem_ctx* c, double* u, double* visc) {
    NEED_INIT;
    if (c->velState != 2)
        return c->fail(LBDEM_E_STATE, "no velocity held: call lbdem_store_velocity before the LB step to read");
    const int64_t N = c->box.g.N;
    if (u) {
        double* d = c->tmp(size_t(3) * N * 8);
        if (!d) return c->fail(LBDEM_E_NOMEM, "device allocation failed");
        lb_download_velocity(c->box, d, c->stream);
        HIPCHK(hipMemcpyAsync(u, d, size_t(3) * N * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (visc && c->box.visc) {
        double* d = c->tmp(size_t(N) * 8);
        if (!d) return c->fail(LBDEM_E_NOMEM, "device allocation failed");
        lb_download_visc(c->box, d, c->stream);
        HIPCHK(hipMemcpyAsync(visc, d, size_t(N) * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else if (visc) {  // every active node relaxes with the run's viscosity
        std::vector<uint8_t> t(N);
        HIPCHK(hipMemcpyAsync(t.data(), c->box.type, N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < N; ++i) {
            const uint8_t b = t[i] & T_BASE;
            visc[i] = (b == T_FLUID || b == T_IF) ? c->k.visc : 0.0;
        }
    }
    return 0;
}

extern "C" int lbdem_get_fields(lbdem_ctx* c, double* f, double* mass, double* n) {
    NEED_INIT;
    const int64_t N = c->box.g.N;
    if (f) {
        double* d = c->tmp(size_t(Q) * N * 8);
        if (!d) return c->fail(LBDEM_E_NOMEM, "device allocation failed");
        lb_download_f(c->box, d, c->stream);
        HIPCHK(hipMemcpyAsync(f, d, size_t(Q) * N * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    std::vector<uint8_t> t(N);
    HIPCHK(hipMemcpyAsync(t.data(), c->box.type, N, hipMemcpyDeviceToHost, c->stream));
    if (mass) {
        if (c->box.prec == LBDEM_PREC_FP64) {
            HIPCHK(hipMemcpyAsync(mass, c->box.mass, N * 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        } else {
            std::vector<float> mf(N);
            HIPCHK(hipMemcpyAsync(mf.data(), c->box.mass, N * 4, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int64_t i = 0; i < N; ++i) mass[i] = mf[i];
        }
        for (int64_t i = 0; i < N; ++i) {
            const uint8_t b = t[i] & T_BASE;
            if (b != T_FLUID && b != T_IF) mass[i] = 0.0;
        }
    }
    if (n) {
        double* d = c->tmp(size_t(N) * 8);
        if (!d) return c->fail(LBDEM_E_NOMEM, "device allocation failed");
        lb_download_n(c->box, c->k, d, c->stream);
        HIPCHK(hipMemcpyAsync(n, d, N * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// lbdem_get_fields through the file-scope scratch, one field (or population plane) at a time:
// the same kernels and values, no allocation on the host or the device. Calls on one GPU share
// its scratch and must not overlap.
// (error messages are string literals: even the error paths allocate nothing)
#define NOALLOC_CHK(x)                                                                        \
    do {                                                                                      \
        if ((x) != hipSuccess) return c->fail_fixed(LBDEM_E_DEVICE, "HIP error at " #x);      \
    } while (0)
extern "C" int lbdem_get_fields_noalloc(lbdem_ctx* c, double* f, double* mass, double* n) {
    if (!c) return LBDEM_E_ARG;
    if (!c->initialised) return c->fail_fixed(LBDEM_E_STATE, "call lbdem_init first");
    const int64_t N = c->box.g.N;
    if (N > int64_t(NOALLOC_DOUBLES))
        return c->fail_fixed(LBDEM_E_CAPACITY,
                             "lbdem_get_fields_noalloc: the box has more than 2097152 (2^21) nodes, the most its "
                             "scratch holds; use lbdem_get_fields");
    double* d = noalloc_scratch(c->prm.device);
    if (!d) return c->fail_fixed(LBDEM_E_STATE, "lbdem_get_fields_noalloc: scratch not set up on this device");
    // stream order: each copy out of the scratch completes before the next kernel overwrites it
    if (f)
        for (int q = 0; q < Q; ++q) {
            lb_download_f(c->box, d, c->stream, q, q + 1);
            NOALLOC_CHK(hipMemcpyAsync(f + q * N, d, size_t(N) * 8, hipMemcpyDeviceToHost, c->stream));
        }
    if (mass) {
        lb_download_mass(c->box, d, c->stream);
        NOALLOC_CHK(hipMemcpyAsync(mass, d, size_t(N) * 8, hipMemcpyDeviceToHost, c->stream));
    }
    if (n) {
        lb_download_n(c->box, c->k, d, c->stream);
        NOALLOC_CHK(hipMemcpyAsync(n, d, size_t(N) * 8, hipMemcpyDeviceToHost, c->stream));
    }
    NOALLOC_CHK(hipGetLastError());
    NOALLOC_CHK(hipStreamSynchronize(c->stream));
    return 0;
}
#undef NOALLOC_CHK

static int fetch(lbdem_ctx* c, double* dst, const double* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, c->stream));
    return 0;
}

extern "C" int lbdem_get_particles(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    const int np = c->dp.nPart;
    if (np == 0) return 0;
    const DemState& d = c->ds;
    const double* srcs[] = {d.x0, d.x1, d.w1, d.FHydro, d.MHydro, d.FContact, d.MContact, d.FLub, d.MLub, d.FRep,
                            d.FTotal, d.MTotal, d.FGrav};
    const int offs[] = {LBDEM_P_X0, LBDEM_P_X1, LBDEM_P_W1, LBDEM_P_FHYDRO, LBDEM_P_MHYDRO, LBDEM_P_FCONTACT,
                        LBDEM_P_MCONTACT, LBDEM_P_FLUB, LBDEM_P_MLUB, LBDEM_P_FREP, LBDEM_P_FTOTAL, LBDEM_P_MTOTAL,
                        LBDEM_P_FGRAV};
    std::vector<double> buf(3 * size_t(np));
    for (int k = 0; k < 13; ++k) {
        if (fetch(c, buf.data(), srcs[k], 3 * np)) return LBDEM_E_DEVICE;
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < np; ++i) for (int a = 0; a < 3; ++a) rec[size_t(i) * LBDEM_P_REC + offs[k] + a] = buf[3 * i + a];
    }
    const double* s2[] = {d.xiw, d.xio};
    const int o2[] = {LBDEM_P_XIW, LBDEM_P_XIO};
    for (int k = 0; k < 2; ++k) {
        if (fetch(c, buf.data(), s2[k], 3 * np)) return LBDEM_E_DEVICE;
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < np; ++i) for (int a = 0; a < 3; ++a) rec[size_t(i) * LBDEM_P_REC + o2[k] + a] = buf[3 * i + a];
    }
    const double* s1[] = {d.fluidVol, d.r, d.m};
    const int o1[] = {LBDEM_P_FLUIDVOL, LBDEM_P_R, LBDEM_P_M};
    for (int k = 0; k < 3; ++k) {
        if (fetch(c, buf.data(), s1[k], np)) return LBDEM_E_DEVICE;
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < np; ++i) rec[size_t(i) * LBDEM_P_REC + o1[k]] = buf[i];
    }
    return 0;
}

extern "C" int lbdem_get_objects(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    const int no = c->dp.nObj;
    if (no == 0) return 0;
    const DemState& d = c->ds;
    const double* srcs[] = {d.ox0, d.ox1, d.ow1, d.oFHydro, d.oMHydro, d.oFContact, d.oMContact, d.oFLub, d.oMLub,
                            d.oFTotal, d.oMTotal, d.oFGrav};
    const int offs[] = {LBDEM_O_X0, LBDEM_O_X1, LBDEM_O_W1, LBDEM_O_FHYDRO, LBDEM_O_MHYDRO, LBDEM_O_FCONTACT,
                        LBDEM_O_MCONTACT, LBDEM_O_FLUB, LBDEM_O_MLUB, LBDEM_O_FTOTAL, LBDEM_O_MTOTAL, LBDEM_O_FGRAV};
    std::vector<double> buf(3 * size_t(no));
    for (int k = 0; k < 12; ++k) {
        if (fetch(c, buf.data(), srcs[k], 3 * no)) return LBDEM_E_DEVICE;
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < no; ++i) for (int a = 0; a < 3; ++a) rec[size_t(i) * LBDEM_O_REC + offs[k] + a] = buf[3 * i + a];
    }
    const double* s1[] = {d.or_, d.om};
    const int o1[] = {LBDEM_O_R, LBDEM_O_M};
    for (int k = 0; k < 2; ++k) {
        if (fetch(c, buf.data(), s1[k], no)) return LBDEM_E_DEVICE;
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < no; ++i) rec[size_t(i) * LBDEM_O_REC + o1[k]] = buf[i];
    }
    return 0;
}

extern "C" int lbdem_get_bodies(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    const int nb = c->dp.nBody;
    if (nb == 0) return 0;
    if (fetch(c, rec, c->ds.body, nb * LBDEM_B_REC)) return LBDEM_E_DEVICE;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int lbdem_set_stress(lbdem_ctx* c, int enable) {
    if (!c) return LBDEM_E_ARG;
    if (c->initialised) return c->fail(LBDEM_E_STATE, "lbdem_set_stress after lbdem_init");
    c->stress = enable != 0;
    return 0;
}

// The per-particle records of the last DEM step, combined in the reference's order on the host:
// STotal = SHydro + SLub + SContact + SRep (DEM.cpp:748), the sums over the particles divided by
// the volume (DEM.cpp:745-771). Each particle pair added its contact and electrostatic tensor to
// both of its spheres, so the pair parts of contactStress and electroStress are half the sums.
extern "C" int lbdem_get_stress(lbdem_ctx* c, double* global, double* stotal, double* disp) {
    NEED_INIT;
    if (!c->stress) return c->fail(LBDEM_E_STATE, "the stress diagnostics are off (lbdem_set_stress before lbdem_init)");
    const int np = c->dp.nPart;
    std::vector<double> rec(size_t(std::max(np, 1)) * SP_REC, 0.0);
    if (np > 0) {
        if (fetch(c, rec.data(), c->ds.stressPart, np * SP_REC)) return LBDEM_E_DEVICE;
        HIPCHK(hipStreamSynchronize(c->stream));
        if (disp) {
            if (fetch(c, disp, c->ds.disp, 3 * np)) return LBDEM_E_DEVICE;
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    }
    double hydro[9] = {}, lub[9] = {}, cpair[9] = {}, cwall[9] = {}, rep[9] = {};
    for (int i = 0; i < np; ++i) {
        const double* R = &rec[size_t(i) * SP_REC];
        for (int e = 0; e < 9; ++e) {
            hydro[e] += R[SP_HYDRO + e]; lub[e] += R[SP_LUB + e];
            cpair[e] += R[SP_CPAIR + e]; cwall[e] += R[SP_CWALL + e]; rep[e] += R[SP_REP + e];
            if (stotal)
                stotal[size_t(i) * 9 + e] = R[SP_HYDRO + e] + R[SP_LUB + e] + (R[SP_CPAIR + e] + R[SP_CWALL + e]) + R[SP_REP + e];
        }
    }
    if (global) {
        const double vol = c->dp.demSize[0] * c->dp.demSize[1] * c->dp.demSize[2];
        for (int e = 0; e < 9; ++e) {
            const double h = hydro[e] / vol, l = lub[e] / vol;
            const double ct = (0.5 * cpair[e] + cwall[e]) / vol, el = 0.5 * rep[e] / vol;
            global[e] = ct + l + el + h;
            global[9 + e] = h; global[18 + e] = l; global[27 + e] = ct; global[36 + e] = el;
        }
    }
    return 0;
}

// STotal as in lbdem_get_stress; SVisc_h = SHydro + SLub and SVisc = its contact part + SVisc_h
// (DEM.cpp:749-750); overlap = FElas / linearStiff (DEM.cpp:761)
extern "C" int lbdem_get_particle_stress(lbdem_ctx* c, double* out) {
    NEED_INIT;
    if (!c->stress) return c->fail(LBDEM_E_STATE, "the stress diagnostics are off (lbdem_set_stress before lbdem_init)");
    const int np = c->dp.nPart;
    if (np == 0) return 0;
    std::vector<double> rec(size_t(np) * SP_REC), disp(3 * size_t(np));
    if (fetch(c, rec.data(), c->ds.stressPart, size_t(np) * SP_REC)) return LBDEM_E_DEVICE;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (fetch(c, disp.data(), c->ds.disp, 3 * size_t(np))) return LBDEM_E_DEVICE;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < np; ++i) {
        const double* R = &rec[size_t(i) * SP_REC];
        double* O = out + size_t(i) * LBDEM_S_REC;
        for (int e = 0; e < 9; ++e) {
            O[LBDEM_S_STOTAL + e] = R[SP_HYDRO + e] + R[SP_LUB + e] + (R[SP_CPAIR + e] + R[SP_CWALL + e]) + R[SP_REP + e];
            O[LBDEM_S_SELAS + e] = R[SP_ELAS + e];
            const double vh = R[SP_HYDRO + e] + R[SP_LUB + e];
            O[LBDEM_S_SVISC_H + e] = vh;
            O[LBDEM_S_SVISC + e] = R[SP_VISC + e] + vh;
        }
        for (int a = 0; a < 3; ++a) {
            O[LBDEM_S_OVERLAP + a] = R[SP_FELAS + a] / c->dp.linearStiff;
            O[LBDEM_S_DISP + a] = disp[3 * size_t(i) + a];
        }
        O[LBDEM_S_COORD] = R[SP_COORD];
        O[LBDEM_S_LUB_COORD] = R[SP_LUBC];
    }
    return 0;
}

extern "C" int lbdem_get_walls(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    const int nw = c->dp.nWall;
    if (nw == 0) return 0;
    const DemState& d = c->ds;
    const double* srcs[] = {d.wFHydro, d.wFContact, d.wFLub, d.wFTotal};
    const int offs[] = {LBDEM_W_FHYDRO, LBDEM_W_FCONTACT, LBDEM_W_FLUB, LBDEM_W_FTOTAL};
    std::vector<double> buf(3 * size_t(nw));
    for (int k = 0; k < 4; ++k) {
        if (fetch(c, buf.data(), srcs[k], 3 * nw)) return LBDEM_E_DEVICE;
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < nw; ++i) for (int a = 0; a < 3; ++a) rec[size_t(i) * LBDEM_W_REC + offs[k] + a] = buf[3 * i + a];
    }
    return 0;
}

extern "C" int lbdem_get_neighbors(lbdem_ctx* c, int cap, int* count, int* list) {
    NEED_INIT;
    const int np = c->dp.nPart;
    if (np == 0) return 0;
    std::vector<int> cnt(np), lst(size_t(np) * NBR_CAP);
    HIPCHK(hipMemcpyAsync(cnt.data(), c->ds.nbrCount, np * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(lst.data(), c->ds.nbrList, size_t(np) * NBR_CAP * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < np; ++i) {
        count[i] = cnt[i];
        for (int k = 0; k < cnt[i] && k < cap; ++k) list[size_t(i) * cap + k] = lst[size_t(i) * NBR_CAP + k];
    }
    return 0;
}

extern "C" int lbdem_total_mass(lbdem_ctx* c, double* m) {
    NEED_INIT;
    const int nb = 1024;
    double* part = c->tmp(nb * sizeof(double));
    if (!part) return c->fail(LBDEM_E_NOMEM, "device allocation failed");
    lb_total_mass(c->box, part, c->stream);
    std::vector<double> h(nb);
    HIPCHK(hipMemcpyAsync(h.data(), part, nb * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : h) s += v;
    *m = s;
    return 0;
}

extern "C" int lbdem_profile_enable(lbdem_ctx* c, int enable) {
    NEED_INIT;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->profiling = enable != 0;
    c->evUsed = 0;
    return 0;
}

extern "C" int lbdem_profile_read(lbdem_ctx* c, double* fused_ms, int* fused_launches) {
    NEED_INIT;
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (size_t i = 0; i + 1 < c->evUsed; i += 2) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->evPool[i], c->evPool[i + 1]));
        tot += ms;
    }
    if (fused_ms) *fused_ms = tot;
    if (fused_launches) *fused_launches = int(c->evUsed / 2);
    return 0;
}

// ---------------------------------------------------------------------------------------------
// batches: every kernel of a step pair launched once for all member boxes
// ---------------------------------------------------------------------------------------------
namespace {
struct BatchOp {
    void (*launch)(dim3, dim3, size_t, hipStream_t, const void*);
    dim3 grid, block;
    size_t shm;
    size_t off;               // of the launch's argument table (one entry per box of the launch)
    size_t j;                 // index of the launch in the recorded step pair
    std::vector<size_t> who;  // the batch members it runs, in table order
    std::string name;         // the kernel body with its template arguments
};
// "k_fused<float, 0, true, true, 4>" from a launcher's demangled type name, whose first template
// argument is the body: "&lbdem::k_mark" or, for a function template,
// "&(void lbdem::k_fused<float, 0, true, true, 4>(lbdem::FusedArgs<float>))"
static std::string body_name(const char* sig) {
    const char* p = sig ? std::strstr(sig, "Launcher<&") : nullptr;
    if (!p) return sig ? sig : "?";
    p += 10;
    if (*p == '(') {  // skip "(void "
        const char* sp = std::strchr(p, ' ');
        p = sp ? sp + 1 : p + 1;
    }
    if (std::strncmp(p, "lbdem::", 7) == 0) p += 7;
    std::string out;
    int depth = 0;
    for (; *p; ++p) {
        if (*p == '<') ++depth;
        else if (*p == '>') {
            if (depth == 0) break;
            --depth;
        } else if ((*p == ',' || *p == '(' || *p == ')') && depth == 0) break;
        out += *p;
    }
    return out;
}
// A kernel with an exact grid runs every box of a launch on the launch's largest grid: the
// surplus blocks of smaller boxes return at once, which costs less than launching boxes of
// different sizes separately (measured on boxes of edge 48-128: 4 % per step).
// the phase's step pair, one launch list per group of members with the same launch sequence
struct BatchPhase {
    bool recorded = false;
    std::vector<std::vector<BatchOp>> groups;
    char* table = nullptr;
    size_t tableBytes = 0;
    bool ownTable = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool captureFailed = false;  // launched op by op from then on
};
}  // namespace

struct lbdem_batch {
    std::vector<lbdem_ctx*> m;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<hipEvent_t> mev;
    BatchPhase ph[4];
    char* region = nullptr;  // caller-owned memory for the argument tables (else hipMalloc)
    size_t regionSize = 0, regionUsed = 0;
    // z-walk of every member; while !tuned the next step pairs run one candidate each, timed (the
    // walks 8, 4 and 2: a batch has blocks enough that the 1-node walk never wins)
    static constexpr int NCAND = 3;
    int tile = 8;
    bool tuned = true;
    int tuneIdx = 0;
    hipEvent_t tuneEv[NCAND][2] = {};
    // kernel timing (lbdem_batch_profile_enable): every launch of a pair bracketed by events
    bool prof = false;
    std::vector<hipEvent_t> pev;
    std::vector<std::string> profName;  // kernels in order of first launch
    std::map<std::string, std::pair<double, int>> profTime;  // ms summed, launches
    std::string err;
    int fail(int code, const std::string& e) { err = e; return code; }
    ~lbdem_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& pr : tuneEv)
            for (hipEvent_t e : pr)
                if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : pev) (void)hipEventDestroy(e);
        for (BatchPhase& P : ph) {
            if (P.exec) (void)hipGraphExecDestroy(P.exec);
            if (P.graph) (void)hipGraphDestroy(P.graph);
            if (P.ownTable && P.table) (void)hipFree(P.table);
        }
        for (hipEvent_t e : mev) (void)hipEventDestroy(e);
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
        for (lbdem_ctx* c : m) c->batch = nullptr;
    }
};

// the contexts' start-up is over: the launches of a step pair repeat (see graph_ready)
static bool steady(const lbdem_ctx* c) {
    if (!c->initialised || c->profiling || !c->tuned) return false;
    if (c->prm.demSolve && c->demTimeStep < unsigned(std::max(c->prm.objWaitTime, 0)) + 1u) return false;
    if (!(c->demTime >= c->prm.demInitialRepeat)) return false;
    if (!c->forceField && !c->lbFReset) return false;
    return true;
}

#define BATCHCHK(x)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return B->fail(LBDEM_E_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

// records one step pair of every member, groups the members whose launch sequences agree
// (kernels, blocks, argument sizes) and uploads the argument tables
static int batch_record(lbdem_batch* B, BatchPhase& P) {
    const size_t n = B->m.size();
    std::vector<Recorder> rec(n);
    std::vector<HostStepState> saved(n);
    for (size_t i = 0; i < n; ++i) {
        lbdem_ctx* c = B->m[i];
        saved[i] = save_host(c);
        current_recorder() = &rec[i];
        enqueue_step(c);
        enqueue_step(c);
        current_recorder() = nullptr;
    }
    auto undo = [&](const std::string& why) {
        for (size_t i = 0; i < n; ++i) restore_host(B->m[i], saved[i]);
        return B->fail(LBDEM_E_STATE, why);
    };
    auto same_sequence = [&](size_t a, size_t b) {
        if (rec[a].ops.size() != rec[b].ops.size()) return false;
        for (size_t j = 0; j < rec[a].ops.size(); ++j) {
            const RecordedLaunch &x = rec[a].ops[j], &y = rec[b].ops[j];
            if (x.key != y.key || x.args.size() != y.args.size() || x.block.x != y.block.x ||
                x.block.y != y.block.y || x.block.z != y.block.z || x.shm != y.shm)
                return false;
        }
        return true;
    };
    std::vector<std::vector<size_t>> members;  // group -> member indices, in member order
    for (size_t i = 0; i < n; ++i) {
        size_t g = 0;
        while (g < members.size() && !same_sequence(members[g][0], i)) ++g;
        if (g == members.size()) members.emplace_back();
        members[g].push_back(i);
    }
    P.groups.assign(members.size(), {});
    size_t total = 0;
    for (size_t g = 0; g < members.size(); ++g) {
        const std::vector<size_t>& mg = members[g];
        const Recorder& r0 = rec[mg[0]];
        for (size_t j = 0; j < r0.ops.size(); ++j) {
            BatchOp op;
            op.launch = r0.ops[j].batch;
            op.block = r0.ops[j].block;
            op.shm = r0.ops[j].shm;
            op.j = j;
            op.name = body_name(r0.ops[j].name);
            unsigned gx = 1;
            for (size_t i : mg) {
                const dim3 gr = rec[i].ops[j].grid;
                if (gr.y != 1 || gr.z != 1) return undo("a kernel uses a multi-dimensional grid");
                gx = std::max(gx, gr.x);
            }
            // grid-stride kernels: a bounded number of blocks spread over the launch's boxes
            if (r0.ops[j].elastic) gx = std::min(gx, std::max(4u, unsigned(r0.ops[j].elastic / mg.size())));
            op.grid = dim3(gx, unsigned(mg.size()), 1);
            op.off = total;
            total += (mg.size() * r0.ops[j].args.size() + 255) & ~size_t(255);
            op.who = mg;
            P.groups[g].push_back(std::move(op));
        }
    }
    std::vector<char> host(std::max<size_t>(total, 1));
    for (const std::vector<BatchOp>& ops : P.groups)
        for (const BatchOp& op : ops) {
            const size_t sz = rec[op.who[0]].ops[op.j].args.size();
            for (size_t k = 0; k < op.who.size(); ++k)
                std::memcpy(host.data() + op.off + k * sz, rec[op.who[k]].ops[op.j].args.data(), sz);
        }
    if (!P.table || P.tableBytes < total) {
        if (P.ownTable && P.table) (void)hipFree(P.table);
        P.table = nullptr;
        P.ownTable = false;
        if (B->region && B->regionUsed + total <= B->regionSize) {
            P.table = B->region + B->regionUsed;
            B->regionUsed += (total + 255) & ~size_t(255);
        } else {
            if (hipMalloc(&P.table, std::max<size_t>(total, 256)) != hipSuccess) {
                P.table = nullptr;
                return undo("device allocation of the batch tables failed");
            }
            P.ownTable = true;
        }
        P.tableBytes = total;
    }
    if (total && (hipMemcpyAsync(P.table, host.data(), total, hipMemcpyHostToDevice, B->stream) != hipSuccess ||
                  hipStreamSynchronize(B->stream) != hipSuccess))
        return undo("upload of the batch tables failed");
    P.recorded = true;
    return 0;
}

static void batch_launch_ops(lbdem_batch* B, const BatchPhase& P) {
    for (const std::vector<BatchOp>& ops : P.groups)
        for (const BatchOp& op : ops) op.launch(op.grid, op.block, op.shm, B->stream, P.table + op.off);
}

// the recorded pair captured into a graph once (the ops are launched directly when capture is
// unavailable)
static void batch_prepare_pair(lbdem_batch* B, BatchPhase& P) {
    if (!P.exec && !P.graph && !P.captureFailed) {
        if (hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            batch_launch_ops(B, P);
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(B->stream, &g) == hipSuccess && g &&
                hipGraphInstantiate(&P.exec, g, nullptr, nullptr, 0) == hipSuccess) {
                P.graph = g;
            } else {
                if (g) (void)hipGraphDestroy(g);
                P.exec = nullptr;
            }
        }
        P.captureFailed = !P.exec;
        (void)hipGetLastError();
    }
}
// profiling: the pair's launches one by one, an event after each, times read back at once (the
// host waits for the pair)
static int batch_launch_pair_timed(lbdem_batch* B, const BatchPhase& P) {
    size_t nops = 0;
    for (const std::vector<BatchOp>& ops : P.groups) nops += ops.size();
    while (B->pev.size() < nops + 1) {
        hipEvent_t e = nullptr;
        BATCHCHK(hipEventCreate(&e));
        B->pev.push_back(e);
    }
    BATCHCHK(hipEventRecord(B->pev[0], B->stream));
    size_t k = 0;
    for (const std::vector<BatchOp>& ops : P.groups)
        for (const BatchOp& op : ops) {
            op.launch(op.grid, op.block, op.shm, B->stream, P.table + op.off);
            BATCHCHK(hipGetLastError());
            BATCHCHK(hipEventRecord(B->pev[++k], B->stream));
        }
    BATCHCHK(hipEventSynchronize(B->pev[nops]));
    k = 0;
    for (const std::vector<BatchOp>& ops : P.groups)
        for (const BatchOp& op : ops) {
            float ms = 0.0f;
            BATCHCHK(hipEventElapsedTime(&ms, B->pev[k], B->pev[k + 1]));
            ++k;
            auto it = B->profTime.find(op.name);
            if (it == B->profTime.end()) {
                B->profName.push_back(op.name);
                it = B->profTime.emplace(op.name, std::make_pair(0.0, 0)).first;
            }
            it->second.first += ms;
            it->second.second += 1;
        }
    return 0;
}
static int batch_launch_pair(lbdem_batch* B, const BatchPhase& P) {
    if (B->prof) return batch_launch_pair_timed(B, P);
    if (P.exec) BATCHCHK(hipGraphLaunch(P.exec, B->stream));
    else {
        batch_launch_ops(B, P);
        BATCHCHK(hipGetLastError());
    }
    return 0;
}

// every member on walk fz: the recorded pairs hold the previous walk and are dropped (after the
// device is done with their tables, which the next recording may overwrite)
static int batch_use_tile(lbdem_batch* B, int fz) {
    bool same = true;
    for (lbdem_ctx* c : B->m) same = same && c->box.fz == fz;
    B->tile = fz;
    if (same) return 0;
    BATCHCHK(hipStreamSynchronize(B->stream));
    for (BatchPhase& P : B->ph) {
        if (P.exec) (void)hipGraphExecDestroy(P.exec);
        if (P.graph) (void)hipGraphDestroy(P.graph);
        if (P.ownTable && P.table) (void)hipFree(P.table);
        P = BatchPhase();
    }
    B->regionUsed = 0;
    for (lbdem_ctx* c : B->m) c->box.fz = fz;
    return 0;
}

extern "C" int lbdem_batch_create(lbdem_ctx* const* ctxs, int n, lbdem_batch** out) {
    if (!ctxs || n < 1 || !out) return LBDEM_E_ARG;
    *out = nullptr;
    lbdem_ctx* c0 = ctxs[0];
    for (int i = 0; i < n; ++i) {
        lbdem_ctx* c = ctxs[i];
        if (!c || !c->initialised) return LBDEM_E_STATE;
        if (c->batch) return c->fail(LBDEM_E_STATE, "the context is already in a batch");
        if (c->stress) return c->fail(LBDEM_E_ARG, "the stress diagnostics are not available in batches");
        if (!steady(c)) return c->fail(LBDEM_E_STATE, "the context has not finished its start-up steps");
        if (c->box.prec != c0->box.prec || c->prm.device != c0->prm.device || c->freeSurface != c0->freeSurface)
            return c->fail(LBDEM_E_ARG, "batch members need the same precision, device and free-surface setting");
        if (graph_phase(c) != graph_phase(c0)) return c->fail(LBDEM_E_STATE, "batch members must be at the same step parity");
        for (int k = 0; k < i; ++k)
            if (ctxs[k] == c) return c->fail(LBDEM_E_ARG, "a context appears twice");
    }
    lbdem_batch* B = new lbdem_batch();
    B->m.assign(ctxs, ctxs + n);
    B->device = c0->prm.device;
    if (hipSetDevice(B->device) != hipSuccess || hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&B->done, hipEventDisableTiming) != hipSuccess) {
        delete B;
        return LBDEM_E_DEVICE;
    }
    B->mev.resize(n, nullptr);
    for (hipEvent_t& e : B->mev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { delete B; return LBDEM_E_DEVICE; }
    // one z-walk tile for all members: the one tuned for the largest box
    const lbdem_ctx* big = c0;
    for (lbdem_ctx* c : B->m) if (c->box.g.N > big->box.g.N) big = c;
    const int fz = big->box.fz;
    B->tile = fz;
    for (lbdem_ctx* c : B->m) {
        (void)hipStreamSynchronize(c->stream);
        drop_graphs(c);  // captured with the context's own tile
        c->box.fz = fz;
        c->batch = B;
        c->velState = 0;  // batched steps keep no velocity
    }
    *out = B;
    return 0;
}

extern "C" int lbdem_batch_cycle(lbdem_batch* B, int nsteps) {
    if (!B) return LBDEM_E_ARG;
    if (nsteps < 0 || nsteps % 2) return B->fail(LBDEM_E_ARG, "a batch advances whole step pairs (nsteps even)");
    if (nsteps == 0) return 0;
    const size_t n = B->m.size();
    for (size_t i = 0; i < n; ++i) {  // after each member's own pending work
        BATCHCHK(hipEventRecord(B->mev[i], B->m[i]->stream));
        BATCHCHK(hipStreamWaitEvent(B->stream, B->mev[i], 0));
    }
    for (int s = 0; s < nsteps; s += 2) {
        int slot = -1;  // tuning: the candidate this pair times
        if (!B->tuned) {
            slot = B->tuneIdx++;
            if (int e = batch_use_tile(B, lbdem_ctx::CAND[slot])) return e;
        }
        BatchPhase& P = B->ph[graph_phase(B->m[0])];
        if (!P.recorded) {
            if (int e = batch_record(B, P)) return e;  // advances the members' counters
        } else {
            for (lbdem_ctx* c : B->m) advance_pair(c);
        }
        batch_prepare_pair(B, P);  // host work done before the timed launch
        if (slot >= 0) BATCHCHK(hipEventRecord(B->tuneEv[slot][0], B->stream));
        if (int e = batch_launch_pair(B, P)) return e;
        if (slot >= 0) {
            BATCHCHK(hipEventRecord(B->tuneEv[slot][1], B->stream));
            if (B->tuneIdx == lbdem_batch::NCAND) {  // every candidate timed: keep the fastest
                BATCHCHK(hipEventSynchronize(B->tuneEv[slot][1]));
                int best = 0;
                float bestMs = 0.0f;
                for (int q = 0; q < lbdem_batch::NCAND; ++q) {
                    float ms = 0.0f;
                    BATCHCHK(hipEventElapsedTime(&ms, B->tuneEv[q][0], B->tuneEv[q][1]));
                    if (q == 0 || ms < bestMs) { bestMs = ms; best = q; }
                }
                B->tuned = true;
                if (int e = batch_use_tile(B, lbdem_ctx::CAND[best])) return e;
            }
        }
    }
    BATCHCHK(hipEventRecord(B->done, B->stream));
    for (lbdem_ctx* c : B->m) BATCHCHK(hipStreamWaitEvent(c->stream, B->done, 0));
    return 0;
}

extern "C" int lbdem_batch_set_tile(lbdem_batch* B, int tileZ) {
    if (!B) return LBDEM_E_ARG;
    if (tileZ != 0 && tileZ != 1 && tileZ != 2 && tileZ != 4 && tileZ != 8)
        return B->fail(LBDEM_E_ARG, "tile z-walk length must be 0 (tune), 1, 2, 4 or 8");
    if (tileZ == 0) {
        for (auto& pr : B->tuneEv)
            for (hipEvent_t& e : pr)
                if (!e) BATCHCHK(hipEventCreate(&e));
        B->tuned = false;
        B->tuneIdx = 0;
        return 0;
    }
    B->tuned = true;
    return batch_use_tile(B, tileZ);
}

extern "C" int lbdem_batch_get_tile(const lbdem_batch* B, int* tileZ, int* tuned) {
    if (!B) return LBDEM_E_ARG;
    if (tileZ) *tileZ = B->tile;
    if (tuned) *tuned = B->tuned ? 1 : 0;
    return 0;
}

extern "C" int lbdem_batch_synchronize(lbdem_batch* B) {
    if (!B) return LBDEM_E_ARG;
    BATCHCHK(hipStreamSynchronize(B->stream));
    for (lbdem_ctx* c : B->m)
        if (int e = check_errors(c)) return B->fail(e, c->err);
    return 0;
}

extern "C" int lbdem_batch_profile_enable(lbdem_batch* B, int enable) {
    if (!B) return LBDEM_E_ARG;
    BATCHCHK(hipStreamSynchronize(B->stream));
    B->prof = enable != 0;
    B->profName.clear();
    B->profTime.clear();
    return 0;
}

extern "C" int lbdem_batch_profile_read(const lbdem_batch* B, int i, const char** kernel, double* ms, int* launches,
                                        int* count) {
    if (!B || i < 0) return LBDEM_E_ARG;
    if (count) *count = int(B->profName.size());
    if (kernel) *kernel = nullptr;
    if (size_t(i) >= B->profName.size()) return 0;
    const auto& v = B->profTime.at(B->profName[size_t(i)]);
    if (kernel) *kernel = B->profName[size_t(i)].c_str();
    if (ms) *ms = v.first;
    if (launches) *launches = v.second;
    return 0;
}

extern "C" const char* lbdem_batch_last_error(const lbdem_batch* B) { return B ? B->err.c_str() : "null batch"; }

extern "C" int lbdem_batch_destroy(lbdem_batch* B) {
    delete B;
    return 0;
}

// run pool hook: the batch's argument tables in caller-owned device memory
int lbdem::batch_use_region(lbdem_batch* B, void* base, size_t bytes) {
    if (!B || !base) return LBDEM_E_ARG;
    // the kernels read their argument records with scalar loads: 256-byte aligned tables
    const uintptr_t a = (reinterpret_cast<uintptr_t>(base) + 255) & ~uintptr_t(255);
    const size_t skip = size_t(a - reinterpret_cast<uintptr_t>(base));
    if (skip >= bytes) return LBDEM_E_ARG;
    B->region = reinterpret_cast<char*>(a);
    B->regionSize = (bytes - skip) & ~size_t(255);
    B->regionUsed = 0;
    return 0;
}
hipStream_t lbdem::batch_stream(const lbdem_batch* B) { return B ? B->stream : nullptr; }
bool lbdem::ctx_steady(const lbdem_ctx* c) { return c && steady(c); }
int lbdem::ctx_phase(const lbdem_ctx* c) { return c ? graph_phase(c) : -1; }
int lbdem::ctx_precision(const lbdem_ctx* c) { return c ? c->box.prec : -1; }
