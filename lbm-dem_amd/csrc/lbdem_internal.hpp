// Internal declarations of the MI355X lbdem library: device data layout, lattice constants,
// neighbour rules and the launcher interface shared by lb.hip, dem.hip and api.hip.
//
// Data layout in HBM (one box, N = X*Y*Z cells, index i = x + X*(y + Y*z)):
//   f      : 19 SoA planes of the population offset g_q = f_q - w_q (FP64 / FP32 / scaled FP16),
//            updated in place with the AA pattern (even steps read/write the cell's own slots,
//            odd steps read/write the neighbours' slots); boundary conditions are resolved by
//            the cell that stores, which writes the reflected value into the slot its own load
//            will read next step.
//   type   : uint8 node type (0 fluid, 2 gas, 3 interface, 4 periodic, 7 wall, 12 object),
//            bit 7 = inside-particle flag
//   solid  : int32 solid index (particle / wall / object index, as the reference's solidIndex)
//   mass   : free-surface mass (fluid cells carry their density)
//   aux    : density, shifted velocity and hydrodynamic force of interface cells (donor data)
// Particles, objects and walls are FP64 (positions, velocities, forces, reductions).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>

#include "../../include/lbdem.h"

namespace lbdem {

constexpr int Q = 19;
constexpr int NBR_CAP = 64;        // Verlet partners per particle
constexpr int MAX_OBJECTS = 8;
constexpr int MAX_WALLS = 6;
constexpr float FP16_SCALE = 4096.0f;   // FP16 storage holds g * 2^12

// D3Q19 (reference lattice.h:33-57, 134, 147-150)
__host__ __device__ constexpr int cxq(int q) {
    constexpr int c[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    return c[q];
}
__host__ __device__ constexpr int cyq(int q) {
    constexpr int c[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0};
    return c[q];
}
__host__ __device__ constexpr int czq(int q) {
    constexpr int c[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    return c[q];
}
__host__ __device__ constexpr int oppq(int q) {
    constexpr int o[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    return o[q];
}
__host__ __device__ constexpr double wq(int q) { return q == 0 ? 12.0 / 36.0 : (q < 7 ? 2.0 / 36.0 : 1.0 / 36.0); }

// type byte: bits 0-3 node type; active nodes: bit 6 'edge' (some wrapped neighbour is not active;
// kept up to date before every lattice step), bit 7 inside-particle flag; wall nodes: bits 4-6 the wall that receives the node's link forces
// (its solid index when that is a wall index, else 0; cleared with the solid index when the
// particle flags are rebuilt)
enum : uint8_t { T_FLUID = 0, T_GAS = 2, T_IF = 3, T_PER = 4, T_WALL = 7, T_OBJ = 12, T_EDGE = 0x40, T_PMASK = 0x80,
                 T_BASE = 0x0F, T_WSHIFT = 4 };
// Boundary-link classes of an edge node, 2 bits per direction j = 1..18 at bit 2(j-1):
// 0 active neighbour (no boundary), 1 gas, 2 wall receiving the node's primary wall index
// (bits 36-38), 3 anything else (object node, another wall, unexpected type: resolved from the
// neighbour's type and solid index).
// boundary-link classes of an edge node: 3 bits per direction 1..18, then the primary wall index.
// OBJ0/OBJ1: object node whose solid index is 0/1; OTHER: anything else (type reloaded in the
// fused kernel)
enum : int { BC_NONE = 0, BC_GAS = 1, BC_WALL = 2, BC_OBJ0 = 3, BC_OBJ1 = 4, BC_OTHER = 7, BC_BITS = 3,
             BC_WALL_SHIFT = 54 };

// Geometry of one box: sizes and which faces wrap (reference LB.cpp:85-126, 632-678:
// an active node wraps along an axis when its face neighbour on that side is periodic).
struct Geom {
    int X, Y, Z;
    int64_t N;
    int per[6];      // boundary[k] == 4
};

__host__ __device__ inline int64_t cidx(const Geom& g, int x, int y, int z) {
    return int64_t(x) + int64_t(g.X) * (int64_t(y) + int64_t(g.Y) * int64_t(z));
}
// neighbour of an active (or gas) node in direction q: wrapped rule
__host__ __device__ inline int64_t nbr_wrapped(const Geom& g, int x, int y, int z, int q) {
    int nx = x + cxq(q), ny = y + cyq(q), nz = z + czq(q);
    if (nx == 0 && g.per[0]) nx = g.X - 2; else if (nx == g.X - 1 && g.per[1]) nx = 1;
    if (ny == 0 && g.per[2]) ny = g.Y - 2; else if (ny == g.Y - 1 && g.per[3]) ny = 1;
    if (nz == 0 && g.per[4]) nz = g.Z - 2; else if (nz == g.Z - 1 && g.per[5]) nz = 1;
    return cidx(g, nx, ny, nz);
}
// neighbour of a node that is neither wall nor active (object nodes): plain index offset
// (reference LB.cpp:570-572), clamped into the box
__host__ __device__ inline int64_t nbr_plain(const Geom& g, int x, int y, int z, int q) {
    int nx = x + cxq(q), ny = y + cyq(q), nz = z + czq(q);
    nx = nx < 0 ? 0 : (nx >= g.X ? g.X - 1 : nx);
    ny = ny < 0 ? 0 : (ny >= g.Y ? g.Y - 1 : ny);
    nz = nz < 0 ? 0 : (nz >= g.Z ? g.Z - 1 : nz);
    return cidx(g, nx, ny, nz);
}

// Particle view used by the lattice kernels (lattice units), refreshed after each DEM step.
struct PartLB {
    double c[3];     // x0 / unit.Length
    double r;        // r / unit.Length
    double v[3];     // x1 / unit.Speed
    double w[3];     // w1 / unit.AngVel
};
struct ObjLB {
    double c[3];     // x0 / unit.Length
    double r;        // r / unit.Length
    double u[3];     // x1 / unit.Speed (object node velocity)
    double pad;
};
static_assert(sizeof(ObjLB) == 8 * sizeof(double), "ObjLB is staged as 8 doubles");

// Scalars shared between kernels (device memory). Doubles first.
struct DevScalars {
    double extraMass;          // streaming extra mass (LB.cpp:1453)
    double surplus;            // free-surface mass surplus (LB.cpp:1644)
    double maxDisp;            // DEM neighbour-list trigger accumulator (DEM.cpp:460-470)
    double obj_F[MAX_OBJECTS][6];   // hydrodynamic F, M per object (lattice units) of the last LB step
    double wall_F[MAX_WALLS][3];    // hydrodynamic F per wall (lattice units, diagnostics)
    double lastExtraMass, lastSurplus;
    int rebuild;               // DEM rebuilt its list this step
    int newNeighborList;       // DEM::newNeighborList
    int builds;                // neighbour list builds
    int nbrOverflow;
    int typeError;             // streaming met an unsupported neighbour type
    int listOverflow;          // a compaction list overflowed
    int nIf;                   // interface list length (LB step)
    int nFilled, nEmptied, nNewIf;
    int nIfFinal;              // interface cells after removeIsolated
    int nObjList;              // object node list length
    int nChanged;              // nodes whose type changed between active and non-active (or among
                               // non-active types) since the last LB step
    int fullMark;              // edge bits and link classes must be recomputed for the whole box
    int kminP;                 // lowest index of a particle-flagged node (-1: none)
    int needScan;              // findNewActive needs the full scan
    int kstar;                 // first particle node still covered
    int pad[1];
};

template <typename Tst> struct StoreT;
template <> struct StoreT<double> {
    using C = double;
    __device__ static double ld(const double* p) { return *p; }
    __device__ static void st(double* p, double v) { *p = v; }
};
template <> struct StoreT<float> {
    using C = float;
    __device__ static float ld(const float* p) { return *p; }
    __device__ static void st(float* p, float v) { *p = v; }
};
template <> struct StoreT<__half> {
    using C = float;
    __device__ static float ld(const __half* p) { return __half2float(*p) * (1.0f / FP16_SCALE); }
    __device__ static void st(__half* p, float v) { *p = __float2half_rn(v * FP16_SCALE); }
};

// The device state of one box.
struct Box {
    Geom g;
    int prec;
    int fz = 8;                // z-walk length of the fused kernel's tiles (1, 2, 4 or 8)
    void* f;                   // 19*N storage elements
    uint8_t* type;
    int32_t* solid;
    void* mass;                // N (double for FP64, float otherwise)
    void* aux;                 // 7*N compute-type values: rho, u(3), h(3) for interface cells
    int32_t* ifList;           // interface list (LB step)
    void* ifNewMass;           // per-list new mass
    int32_t* ifList2;          // interface list after conversions
    int32_t* filled;
    int32_t* emptied;
    int32_t* newIf;
    int32_t* donor;            // N, -1 = none
    int32_t* flagNew;          // N, -1 = none (particle flags computed at a rebuild)
    int32_t* objList;          // object nodes, ascending
    int32_t* objCand;          // uncovered object nodes of the current step, ascending
    int* coverCount;           // per-block counts of the object cover and uncover passes
    int32_t* changed;          // nodes to re-mark (see DevScalars::nChanged)
    int changedCap;
    int ifCap, objCap;
    double* blockPart;         // per-block partials of the fused kernel (object F/M, extra mass, walls)
    int32_t* blockFlag;        // per block: partials written
    uint64_t* bcm;             // N, boundary-link classes of edge nodes (valid where T_EDGE is set)
    int32_t* slipList;         // nodes with a diagonal link into a wall, when the walls slip (else null)
    int nSlip;
    void* visc;                // N compute-type node viscosities (lattice units), or null: every node
                               // has the run's viscosity (Newtonian, no turbulence, no restart values)
    int blockPartCap;
    double* hacc;              // per-particle hydro accumulators: F3 M3 vol (lattice units)
    int parity;                // 0: populations in natural layout, 1: swapped
};

struct Consts {
    // lattice-unit parameters
    double lbF[3];              // force in lattice units (zero once collision() ran without forceField)
    double initVel[3];
    double visc;                // lattice kinematic viscosity (initDynVisc)
    double rho0;
    double omega, omegaf;       // 1/(0.5+3 nu), 1 - 1/(1+6 nu)
    // apparent viscosity (node.cpp:113-168; with per-node viscosities only): rheology model
    // (LBDEM_RHEO_*), its parameters in lattice units, the Smagorinsky constant, the clamp range
    int rheo, turb, viscUpdate;
    int slip;                   // walls slip (1: boundary0 = 5, 2: boundary0 = 6, moving with zero velocity):
                                // no wall forces; the slip pass follows the fused kernel; 2 also
                                // enters the free-surface mass exchange
    double slipCoef;            // slipCoefficient (LB.cpp:313, 1344-1345)
    double plasticVisc, yieldStress, frictionCoef, voellmyCoef, turbConst, minVisc, maxVisc;
    int freeSurface;
    int nPart, nObj, nWall;
    double unitLength, unitSpeed, unitAngVel, unitForce, unitTorque, unitVolume;
};

// ---------------- launchers (lb.hip) ----------------
void lb_init_fields(const Box& b, const Consts& k, hipStream_t s);
// rec: LBDEM_RESTART_REC values per record, written at active node node[r]
void lb_restart_fields(const Box& b, const double* rec, const int32_t* node, int nrec, hipStream_t s);
// 3 N values: the velocity held in aux for every active node, zero elsewhere
void lb_download_velocity(const Box& b, double* out, hipStream_t s);
// N values: the node viscosity plane (b.visc must be set) at active nodes, zero elsewhere
void lb_download_visc(const Box& b, double* out, hipStream_t s);
// SHydro of every particle (9 per particle, physical units: L, F = unit.Length, unit.Force) from
// the aux planes an LB step with storeVel left
void lb_hydro_stress(const Box& b, const PartLB* plb, int nPart, double L, double F, double* shydro, hipStream_t s);
void lb_mark_edges(const Box& b, DevScalars* sc, hipStream_t s);
// storeVel: every active node's density and velocity into aux (else the interface nodes' only)
void lb_fused_step(Box& b, const Consts& k, const PartLB* plb, const ObjLB* olb, DevScalars* sc, bool diag,
                   bool storeVel, hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void lb_hydro_finalize(const Box& b, const Consts& k, double* FHydro, double* MHydro, double* fluidVol,
                       DevScalars* sc, double* objFH, double* objMH, double* wallFH, bool diag, hipStream_t s);
void lb_free_surface(Box& b, const Consts& k, DevScalars* sc, hipStream_t s);
// particle flags of a neighbour-list rebuild; with useList the Verlet lists of that rebuild
// resolve overlapping spheres (highest index wins) without atomics
// order: the particles sorted by link cell (the rebuild's cell list), or null
void lb_flags_rebuild_compute(const Box& b, const Consts& k, const PartLB* plb, DevScalars* sc, const int* nbrCount,
                              const int* nbrList, bool useList, const int* order, hipStream_t s);
constexpr int PART_STRIDE_ = 32;
void lb_flags_rebuild_apply(const Box& b, DevScalars* sc, hipStream_t s);
void lb_find_new_active(const Box& b, const Consts& k, const PartLB* plb, DevScalars* sc, hipStream_t s);
void lb_object_update(Box& b, const Consts& k, const ObjLB* olb, DevScalars* sc, hipStream_t s);
int fused_grid_blocks(const Geom& g, int fz, int prec);  // blocks of the fused kernel for z-walk length fz
constexpr int FUSED_TILE_X = 64, FUSED_TILE_Y = 4;  // threads of a fused-kernel tile (x, y)
constexpr int COVER_BLOCKS = 512, UNCOVER_BLOCKS_MAX = 1024;  // grids of the impactor node passes
void lb_download_f(const Box& b, double* out, hipStream_t s, int q0 = 0, int q1 = Q);
void lb_download_mass(const Box& b, double* out, hipStream_t s);
void lb_download_n(const Box& b, const Consts& k, double* out, hipStream_t s);
void lb_total_mass(const Box& b, double* out, hipStream_t s);
void lb_count_types(const Box& b, int* counts, hipStream_t s);

// ---------------- DEM (dem.hip) ----------------
struct DemParams {
    double demSize[3];
    double deltat;
    double linearStiff, dampCoeff, frictionCoefPart, kinVisc;
    double lubNormal, lubTangential, lubRotation, lubCutoff;
    double electroForce, debyeLength;
    double rmax;
    int objWaitTime;
    int periodic[3];           // both faces periodic along axis
    int nlx, nly, nlz;         // cell grid: all 0 for the direct sum; else nlx, nly >= 3, nlz >= 0
    int nPart, nObj, nWall;
    int nBody, nonLinType;     // body-spring impactor (DEM.cpp:63-73, 798-869)
    double springDamp, epsilon, k13;
    double wallP[MAX_WALLS][3], wallN[MAX_WALLS][3];
    double unitLength, unitSpeed, unitAngVel;
    int stress;                // accumulate the stress diagnostics (DemState::stressPart, disp)
};
// per-particle stress record of the last DEM step (doubles): the stresslet SHydro the force kernel
// copied from DemState::shydro, SLub, the contact stress of the particle pairs (each pair adds the
// same tensor to both of its spheres), the contact stress of walls and impactors, SRep, the
// elastic and viscous parts of the contact stress (SElas, SVisc before SVisc_h is added), FElas,
// the contact and normal-lubrication counts; tensors m00 m01 m02 m10 ... m22
enum { SP_HYDRO = 0, SP_LUB = 9, SP_CPAIR = 18, SP_CWALL = 27, SP_REP = 36, SP_ELAS = 45, SP_VISC = 54,
       SP_FELAS = 63, SP_COORD = 66, SP_LUBC = 67, SP_REC = 68 };
constexpr int VEL_SLOTS = 64;  // slots of the particle speed maximum (DemState::blockMax)
struct DemState {
    // particles (FP64, vectors stored as 3 consecutive doubles)
    double *x0, *x1, *w1, *r, *m, *I, *im, *iI;
    double *FHydro, *MHydro, *FContact, *MContact, *FLub, *MLub, *FRep, *FTotal, *MTotal, *FGrav, *fluidVol;
    double *xiw, *xio;
    int* nbrCount;             // partners per particle
    int* nbrList;              // NBR_CAP per particle, ascending
    int* nbrSlot;              // history slot of each entry (in the list of the higher index)
    int* nbrCountOld;
    int* nbrListOld;
    double* xiCur;             // tangential history per slot (3 per slot), read by the force kernel
    double* xiNext;            // written by the force kernel (buffers swap every step)
    int* cellCount;
    int* cellStart;
    int* cellFill;
    int* cellItems;
    int* cellKey;
    double* objContrib;        // per particle, per object: Fc(3) Mc(3) Fl(3) Ml(3)
    double* wallContrib;       // per particle, per wall: Fc(3) Fl(3)
    double* objSum;            // 12*MAX_OBJECTS
    double* redPart;           // 2*64*12*MAX_OBJECTS block partials of the object/wall reduction
    double* wallSum;           // 6*MAX_WALLS
    unsigned long long* blockMax;  // [VEL_SLOTS]: largest particle |v|^2 of the last integration, per slot (bit patterns)
    // objects
    double *ox0, *ox1, *ow1, *or_, *om, *oI, *oim, *oiI;
    double *oFHydro, *oMHydro, *oFContact, *oMContact, *oFLub, *oMLub, *oFTotal, *oMTotal, *oFGrav;
    // bodies: LBDEM_B_REC doubles each (include/lbdem.h layout)
    double* body;
    // stress diagnostics (null unless DemParams::stress)
    double* stressPart;        // SP_REC per particle
    double* shydro;            // 9 per particle: SHydro of the last LB step (physical units)
    double* xi0;               // 3 per particle: x_i, the position at the start or at objWaitTime
    double* disp;              // 3 per particle: x_i - x0 after the last integration
    // walls
    double *wFHydro, *wFContact, *wFLub, *wFTotal;
    PartLB* plb;
    ObjLB* olb;
};
void dem_neighbor_list(const DemParams& p, DemState& d, DevScalars* sc, bool force, hipStream_t s);
// hydroReady (may be null): an event the integration waits for (the hydrodynamic sums it reads);
// false if the wait could not be enqueued
bool dem_step(const DemParams& p, DemState& d, DevScalars* sc, int demTimeStep, hipStream_t s,
              hipEvent_t hydroReady = nullptr);
void dem_refresh_views(const DemParams& p, DemState& d, hipStream_t s);

// run pool hooks: a context carves its buffers from [base, base+bytes) instead of allocating
// them (set before lbdem_init); the stream a context enqueues on
int ctx_use_region(lbdem_ctx* c, void* base, size_t bytes);
hipStream_t ctx_stream(const lbdem_ctx* c);
// a batch's argument tables in [base, base+bytes) (set before its first cycle); its stream
int batch_use_region(lbdem_batch* b, void* base, size_t bytes);
hipStream_t batch_stream(const lbdem_batch* b);
// start-up over (a batch may take the context); AA/history phase; storage precision
bool ctx_steady(const lbdem_ctx* c);
int ctx_phase(const lbdem_ctx* c);
int ctx_precision(const lbdem_ctx* c);

}  // namespace lbdem
