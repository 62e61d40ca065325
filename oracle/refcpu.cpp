// ORACLE — test infrastructure only.
//
// FP64 CPU restatement of the reference's coupled free-surface LB / DEM timestep, written
// from scratch to the reference's semantics (including its order-dependent rules), behind
// the same C ABI as the MI355X library (include/lbdem.h). Only tests/, the smoke check and
// bench.py's cpu_baseline leg load this library; the product never does.
//
// Every function cites the reference function it restates (paths relative to the
// reference repository). Quantities used only by output-only diagnostics (stress tensors,
// FElas/FVisc, contact-point lists, strains) are not computed.
//
// Pinned behaviour where the reference is undefined (each also used by the GPU library):
//  * xPbcs/yPbcs/zPbcs loop index starts at 0 (DEM.cpp:2771, 2792, 2817 leave it uninitialised).
//  * object-particle tangential damping dampCoeffTanw = 0 (DEM.cpp:916 declares it without an
//    initialiser and never assigns it; the reference program of oracle/Makefile is compiled with
//    every such local zero-initialised, which gives the same value).
//  * wall bounce-back force goes to walls[solidIndex] when solidIndex < nWalls, else walls[0]
//    (LB.cpp:1425-1429 would index walls[] with a particle index).
//  * object bounce-back force and torque are dropped when the object node's solidIndex is not an
//    object index (LB.cpp:1441-1449 indexes objects[] with the particle index that
//    updateIndices, LB.cpp:1875-1899, gives an object node that was active on the previous list);
//    the extra mass of those links is kept (it does not depend on the index).
//  * a particle-flagged node that has no node object (gas) takes the active-node neighbour
//    rule in findNewSolid (LB.cpp:2046 dereferences a null node there).
//  * neighbour capacity is 63 partners per particle (reference: 49, unchecked, elmt.h:44-51);
//    exceeding it is reported as LBDEM_E_CAPACITY.
#include "refcpu_common.h"
#include "../include/lbdem.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace refcpu;

namespace {

constexpr int MAXL = 64;  // list[0] holds the count, entries 1..63

struct Units {  // measureUnits (node.h:192-225, node.cpp:512-525)
    double Length = 1, Time = 1, Density = 1, Volume = 1, Speed = 1, Accel = 1, AngVel = 1, KinVisc = 1,
           DynVisc = 1, Force = 1, Torque = 1, Mass = 1, Stress = 1, Pressure = 1, invLength = 1;
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
        Pressure = Density * Length * Length / Time / Time;
        invLength = 1.0 / Length;
    }
};

struct T9 {  // tinyMatrix (myvector.h), m00 m01 m02 m10 ... m22
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    void reset() { for (double& v : m) v = 0.0; }
    // += -a (x) f, the stress of a force f at arm a (DEM.cpp:1097-1105 and the other pair loops)
    void addArm(double ax, double ay, double az, double fx, double fy, double fz) {
        m[0] += -ax * fx; m[1] += -ax * fy; m[2] += -ax * fz;
        m[3] += -ay * fx; m[4] += -ay * fy; m[5] += -ay * fz;
        m[6] += -az * fx; m[7] += -az * fy; m[8] += -az * fz;
    }
    // += a symmetric stresslet given as xx yy zz yz zx xy (DEM.cpp:1325-1333)
    void addSym(const double* s) {
        m[0] += s[0]; m[1] += s[5]; m[2] += s[4];
        m[3] += s[5]; m[4] += s[1]; m[5] += s[3];
        m[6] += s[4]; m[7] += s[3]; m[8] += s[2];
    }
};

struct Particle {  // particle (elmt.h:15-158), fields used on the path
    unsigned size = 1, particleIndex = 0;
    double r = 1, m = 1, I = 1, _m = 1, _I = 1, dens = 1;
    V3 x0, x1, w1, x0_lst;
    V3 x_i, disp;                             // position at the start (or at objWaitTime) and x_i - x0
    T9 SHydro, SLub, SContact, SRep, STotal;  // stress diagnostics (DEM.cpp:655-771)
    T9 SElas, SVisc, SVisc_h;                 // elastic and viscous parts of the contact stress
    V3 FElas, overlap;
    double coord = 0, lub_coord = 0;          // contacts and normal-lubrication pairs
    V3 FHydro, FContact, FTotal, FLub, FGrav, FRep, MHydro, MContact, MTotal, MLub;
    double fluidVolume = 0;
    int list[MAXL], _list[MAXL];
    double xix[MAXL], xiy[MAXL], xiz[MAXL], _xix[MAXL], _xiy[MAXL], _xiz[MAXL];
    V3 xiw, xio;
    Particle() {
        std::memset(list, 0, sizeof(list));
        std::memset(_list, 0, sizeof(_list));
        for (int i = 0; i < MAXL; ++i) xix[i] = xiy[i] = xiz[i] = _xix[i] = _xiy[i] = _xiz[i] = 0.0;
    }
};

struct Object {  // object (elmt.h:160-228)
    unsigned index = 0;
    double r = 0, m = 1, _m = 1, I = 1, _I = 1, rho = 0;
    V3 x0, x1, w1;
    V3 FContact, FHydro, FLub, FGrav, FSpring, FTotal, MContact, MHydro, MTotal, MLub;
};

struct Body {  // body (elmt.h), spring-connected to the impactors (DEM.cpp:168-196, 255-266)
    double r = 0, m = 1, _m = 1, k = 0, L0 = 0, L = 0;
    V3 x0, x1, FGrav, FSpring, FTotal;
};

struct Wall {  // wall (utils.h:19-101)
    unsigned index = 0;
    V3 n = V3(1, 0, 0), p;
    V3 FHydro, FContact, FLub, FTotal, FRep;
};

}  // namespace

struct lbdem_ctx {
    lbdem_params prm;
    std::string err;
    bool initialised = false;
    // ---------------- LB (LB.h:28-150) ----------------
    Units unit;
    unsigned lbSize[3] = {1, 1, 1};
    double lbPhysicalSize[3];
    V3 lbBoundaryLocation[6];
    int boundary[6];
    int wallType = 7;  // node type of every wall: 7 static, 8 moving (zero velocity), 5 static slip, 6 moving slip
    double slipCoefficient = 0.0;
    double fluidHeight = 0;
    double initDynVisc = 0, initDensity = 1.0;
    // FluidMaterial (node.h:17-49): rheology model and its parameters in lattice units
    int rheologyModel = LBDEM_RHEO_NEWTONIAN;
    double plasticVisc = 0, yieldStress = 0, frictionCoefFluid = 0, voellmyCoefficient = 0, turbConst = 0;
    bool turbulenceOn = false;
    V3 initVelocity, lbF;
    bool freeSurface = false, forceField = false;
    unsigned lbTime = 0;
    int64_t shift[3], domain[3], ne[Q];
    int64_t N = 0;
    double totalMass = 0;
    // node types (node.h:116-175)
    std::vector<uint8_t> t, p;
    std::vector<uint16_t> solid;
    // node objects (node.h:54-114) held as SoA; exists[i] == (nodes[i] != 0)
    std::vector<uint8_t> exists;
    std::vector<double> f, fs, n, u, h, mass, newMass, visc;
    std::vector<uint32_t> d;
    std::vector<uint32_t> activeNodes, fluidNodes, interfaceNodes, particleNodes, objectNodes, wallNodes;
    double lastExtraMass = 0, lastSurplus = 0;
    // ---------------- DEM (DEM.h:22-118) ----------------
    double demSize[3] = {1, 1, 1};
    double vol = 1;  // volume the stresses are divided by (DEM.cpp:212-216)
    T9 totalStress, contactStress, lubStress, electroStress, hydroStress;  // DEM.cpp:655-771
    std::vector<Particle> particles;
    std::vector<Object> objects;
    std::vector<Body> bodies;
    std::vector<Wall> walls;
    double demInitialRepeat = 0;
    unsigned demTimeStep = 0;
    double deltat = 1.0, demTime = 0, maxDisp = 0;
    V3 demF;
    double dampCoeff = 0;
    double minPartRadius = 0, maxPartRadius = 0, meanPartRadius = 0;
    bool newNeighborList = false;
    int nlBuilds = 0;
    // ---------------- IO counters (main.cpp:40-43) ----------------
    double realTime = 0;
    unsigned currentTimeStep = 0;
    // ---------------- fluid restart (lbdem_set_restart_fluid) ----------------
    bool lbRestart = false, restartShort = false;
    std::vector<uint8_t> restartTypes;
    std::vector<double> restartRecords;
    int fail(int code, const std::string& m) { err = m; return code; }

    // ================= index helpers (LB.cpp:2113-2169) =================
    unsigned getX(int64_t i) const { div_t a = div(int(i), int(lbSize[0] * lbSize[1])); div_t b = div(a.rem, int(lbSize[0])); return b.rem; }
    unsigned getY(int64_t i) const { div_t a = div(int(i), int(lbSize[0] * lbSize[1])); div_t b = div(a.rem, int(lbSize[0])); return b.quot; }
    unsigned getZ(int64_t i) const { div_t a = div(int(i), int(lbSize[0] * lbSize[1])); return a.quot; }
    V3 getPosition(int64_t i) const { return V3(double(getX(i)), double(getY(i)), double(getZ(i))); }

    bool isActive(int64_t i) const { return t[i] == 0 || t[i] == 3; }
    bool isWall(int64_t i) const { return t[i] >= 5 && t[i] <= 11; }

    // ================= node methods (node.cpp) =================
    double* F(int64_t i) { return &f[size_t(i) * Q]; }
    double* FS(int64_t i) { return &fs[size_t(i) * Q]; }
    V3 U(int64_t i) const { return V3(u[3 * i], u[3 * i + 1], u[3 * i + 2]); }
    void setU(int64_t i, const V3& v) { u[3 * i] = v.x; u[3 * i + 1] = v.y; u[3 * i + 2] = v.z; }
    V3 H(int64_t i) const { return V3(h[3 * i], h[3 * i + 1], h[3 * i + 2]); }
    void setH(int64_t i, const V3& v) { h[3 * i] = v.x; h[3 * i + 1] = v.y; h[3 * i + 2] = v.z; }

    void equilibrium(double nn, const V3& uu, double* feq) const {  // node.cpp:100-111
        const double usq = uu.norm2();
        for (int j = 0; j < Q; ++j) {
            const double vu = uu.dot(vdir(j));
            feq[j] = coeff(j) * nn * (1.0 + 3.0 * vu + 4.5 * vu * vu - 1.5 * usq);
        }
    }
    // node::initialize (node.cpp:26-37) + setEquilibrium (node.cpp:52-73)
    void nodeInitialize(int64_t i, double density, const V3& velocity, double massFunction, double viscosity, const V3& F_) {
        n[i] = density;
        mass[i] = massFunction;
        V3 uu = F_ / 2.0 / n[i] + velocity;
        setU(i, uu);
        visc[i] = viscosity;
        double feq[Q];
        equilibrium(n[i], uu, feq);
        for (int j = 0; j < Q; ++j) F(i)[j] = FS(i)[j] = feq[j];
    }
    // LB::createNode (LB.cpp:474-481): fresh node (node() constructor, node.h:75-88) + findNeighbors
    void createNode(int64_t it) {
        exists[it] = 1;
        n[it] = 0; mass[it] = newMass[it] = 0; visc[it] = 0;
        setU(it, V3()); setH(it, V3());
        for (int j = 0; j < Q; ++j) F(it)[j] = FS(it)[j] = 0.0;
        findNeighbors(it, &d[size_t(it) * Q]);
    }
    void deleteNode(int64_t it) { exists[it] = 0; }

    // LB::findNeighbors (LB.cpp:567-680)
    void findNeighbors(int64_t it, uint32_t* dd) const {
        int64_t tmp[Q];
        for (int j = 0; j < Q; ++j) tmp[j] = it + ne[j];
        if (isWall(it)) {
            const unsigned xh = getX(it), yh = getY(it), zh = getZ(it);
            for (int j = 1; j < Q; ++j) {
                if (xh == 0 && CX[j] < 0) tmp[j] = it;
                if (xh == lbSize[0] - 1 && CX[j] > 0) tmp[j] = it;
                if (yh == 0 && CY[j] < 0) tmp[j] = it;
                if (yh == lbSize[1] - 1 && CY[j] > 0) tmp[j] = it;
                if (zh == 0 && CZ[j] < 0) tmp[j] = it;
                if (zh == lbSize[2] - 1 && CZ[j] > 0) tmp[j] = it;
            }
        } else if (isActive(it)) {
            int64_t pbc[Q] = {0};
            auto per = [&](int64_t k) { return k >= 0 && k < N && t[k] == 4; };
            if (per(tmp[1])) for (int j = 1; j < Q; ++j) if (CX[j] > 0) pbc[j] -= domain[0];
            if (per(tmp[2])) for (int j = 1; j < Q; ++j) if (CX[j] < 0) pbc[j] += domain[0];
            if (per(tmp[3])) for (int j = 1; j < Q; ++j) if (CY[j] > 0) pbc[j] -= domain[1];
            if (per(tmp[4])) for (int j = 1; j < Q; ++j) if (CY[j] < 0) pbc[j] += domain[1];
            if (per(tmp[5])) for (int j = 1; j < Q; ++j) if (CZ[j] > 0) pbc[j] -= domain[2];
            if (per(tmp[6])) for (int j = 1; j < Q; ++j) if (CZ[j] < 0) pbc[j] += domain[2];
            for (int j = 1; j < Q; ++j) tmp[j] += pbc[j];
        }
        for (int j = 0; j < Q; ++j) dd[j] = uint32_t(tmp[j]);
    }
    // neighbour of a node: the stored link when the node object exists, the active-node
    // rule otherwise (pinned, see header)
    uint32_t link(int64_t i, int j) {
        if (exists[i]) return d[size_t(i) * Q + j];
        uint8_t save = t[i];
        uint32_t dd[Q];
        if (!isActive(i) && !isWall(i)) t[i] = 0;
        findNeighbors(i, dd);
        t[i] = save;
        return dd[j];
    }
    bool inRange(uint32_t k) const { return int64_t(k) < N; }

    // ================= parameters (LB.cpp:128-338, DEM.cpp:17-75) =================
    int getParams() {
        const lbdem_params& P = prm;
        unit.Length = P.unitLength; unit.Time = P.unitTime; unit.Density = P.unitDensity;
        if (!(unit.Length > 0) || !(unit.Time > 0) || !(unit.Density > 0)) return fail(LBDEM_E_ARG, "unit lengths/times must be > 0");
        for (int a = 0; a < 3; ++a) {
            if (!(P.lbSize[a] > 0)) return fail(LBDEM_E_ARG, "lbSize must be > 0");
            lbSize[a] = unsigned(int(std::floor(P.lbSize[a] / unit.Length + 0.5)));
            lbPhysicalSize[a] = double(lbSize[a]) * unit.Length;
        }
        lbBoundaryLocation[0] = V3(0.5 * unit.Length, 0.0, 0.0);
        lbBoundaryLocation[1] = V3(double(lbSize[0] - 1.5) * unit.Length, 0.0, 0.0);
        lbBoundaryLocation[2] = V3(0.0, 0.5 * unit.Length, 0.0);
        lbBoundaryLocation[3] = V3(0.0, double(lbSize[1] - 1.5) * unit.Length, 0.0);
        lbBoundaryLocation[4] = V3(0.0, 0.0, 0.5 * unit.Length);
        lbBoundaryLocation[5] = V3(0.0, 0.0, double(lbSize[2] - 1.5) * unit.Length);
        initDynVisc = P.initVisc;
        if (!(initDynVisc > 0)) return fail(LBDEM_E_ARG, "initVisc must be > 0");
        unit.setComposite();
        initVelocity = V3(P.initVelocity[0], P.initVelocity[1], P.initVelocity[2]);
        initVelocity /= unit.Speed;
        lbF = V3(P.lbF[0], P.lbF[1], P.lbF[2]);
        lbF /= unit.Accel;
        for (int i = 0; i < 6; ++i) {
            boundary[i] = P.boundary[i];
            if (boundary[i] != 4 && (boundary[i] < 5 || boundary[i] > 8))
                return fail(LBDEM_E_ARG, "boundary codes must be 4 (periodic) or 5-8 (walls)");
        }
        fluidHeight = P.fluidHeight;
        rheologyModel = P.rheologyModel;
        if (rheologyModel < LBDEM_RHEO_NEWTONIAN || rheologyModel > LBDEM_RHEO_VOELLMY)
            return fail(LBDEM_E_ARG, "unknown rheologyModel");
        plasticVisc = P.plasticVisc;
        yieldStress = P.yieldStress;
        frictionCoefFluid = P.frictionCoefFluid;
        voellmyCoefficient = P.voellmyConstant;
        turbulenceOn = P.turbulenceSolve != 0;
        turbConst = P.turbConst;
        initDynVisc /= unit.DynVisc;  // LB.cpp:320-323
        plasticVisc /= unit.DynVisc;
        yieldStress /= unit.Stress;
        voellmyCoefficient /= unit.Accel;
        initDensity = 1.0;
        freeSurface = P.freeSurface != 0;
        forceField = P.forceField != 0;
        demInitialRepeat = P.demInitialRepeat;
        // every wall takes boundary0's kind (DEM.cpp:430-446 test externalBoundary[0]); slip walls
        // (5, 6) are not restated
        wallType = (boundary[0] >= 5 && boundary[0] <= 8) ? boundary[0] : 7;
        slipCoefficient = P.slipCoefficient;
        if (!P.lbSolve) return fail(LBDEM_E_ARG, "lbSolve=0 is not supported");
        if (!(P.density > 0) || !(P.objectDensity > 0)) return fail(LBDEM_E_ARG, "density/objectDensity must be > 0");
        return 0;
    }

    // ================= LB::latticeDefinition (LB.cpp:85-126) =================
    void latticeDefinition() {
        shift[2] = int64_t(lbSize[0]) * lbSize[1];
        shift[1] = lbSize[0];
        shift[0] = 1;
        domain[2] = int64_t(lbSize[2]) * lbSize[1] * lbSize[0] - 2 * shift[2];
        domain[1] = int64_t(lbSize[1]) * lbSize[0] - 2 * shift[1];
        domain[0] = int64_t(lbSize[0]) - 2 * shift[0];
        for (int j = 0; j < Q; ++j) ne[j] = CX[j] * shift[0] + CY[j] * shift[1] + CZ[j] * shift[2];
    }

    // ================= DEM init (DEM.cpp:199-357, elmt.cpp:17-53, DEM.cpp:405-458) =================
    void discreteElementInit() {
        demSize[0] = lbPhysicalSize[0];
        demSize[1] = freeSurface ? fluidHeight : lbPhysicalSize[1];
        demSize[2] = lbPhysicalSize[2];
        vol = freeSurface ? demSize[0] * fluidHeight * demSize[2] : demSize[0] * demSize[1] * demSize[2];
        const double partDensity = prm.density, objDensity = prm.objectDensity;
        demF = lbF * unit.Accel;
        unsigned globalIndex = 0;
        for (auto& P : particles) {
            const double singleMass = 4.0 / 3.0 * partDensity * M_PI * P.r * P.r * P.r;
            P.m = P.size * singleMass;
            P.dens = partDensity;
            P.I = P.size * 2.0 / 5.0 * singleMass * P.r * P.r;
            P.x1.reset(); P.w1.reset();
            P.particleIndex = globalIndex;
            P.FTotal.reset(); P.FRep.reset(); P.FHydro.reset(); P.FContact.reset(); P.FLub.reset();
            P.FGrav = demF * P.m;
            P.MTotal.reset(); P.MHydro.reset(); P.MContact.reset(); P.MLub.reset();
            P.x_i = P.x0;
            globalIndex++;
        }
        for (auto& O : objects) {
            O.m = 4.0 / 3.0 * objDensity * M_PI * O.r * O.r * O.r;
            O.FGrav = demF * O.m;
            O.I = 2.0 / 5.0 * O.m * O.r * O.r;
            O._m = 1.0 / O.m;
            O._I = 1.0 / O.I;
            O.rho = objDensity;
        }
        for (auto& B : bodies) {  // DEM.cpp:257-266
            B.m = prm.bodyDensity * (4.0 / 3.0) * M_PI * (B.r * B.r * B.r);
            B._m = 1.0 / B.m;
            B.FGrav = demF * B.m;
            B.L = B.L0;
        }
        // DEM::initializeWalls (DEM.cpp:405-458); walls are static no-slip (type 7 only)
        walls.clear();
        const V3 dir[6] = {V3(1, 0, 0), V3(-1, 0, 0), V3(0, 1, 0), V3(0, -1, 0), V3(0, 0, 1), V3(0, 0, -1)};
        unsigned index = 0;
        for (int i = 0; i < 6; ++i) {
            if (boundary[i] >= 5 && boundary[i] <= 8) {
                Wall w; w.p = lbBoundaryLocation[i]; w.n = dir[i]; w.index = index++;
                walls.push_back(w);
            }
        }
        for (auto& P : particles) { P._m = 1.0 / P.m; P._I = 1.0 / P.I; }
        neighborList();
        newNeighborList = true;
        minPartRadius = 1.0e99; maxPartRadius = 0.0; meanPartRadius = 0.0;
        for (auto& P : particles) {
            meanPartRadius += P.r;
            if (P.r > maxPartRadius) maxPartRadius = P.r;
            if (P.r < minPartRadius) minPartRadius = P.r;
        }
        if (particles.size() > 0) meanPartRadius /= double(particles.size());
        deltat = unit.Time / double(1);
        const double averageMass = 4.0 / 3.0 * partDensity * M_PI * meanPartRadius * meanPartRadius * meanPartRadius;
        dampCoeff = std::sqrt(averageMass * prm.linearStiff);
    }

    // ================= LB init (LB.cpp:340-398 and helpers) =================
    void latticeBolzmannInit() {
        // initializeNodes (LB.cpp:500-536)
        N = int64_t(lbSize[0]) * lbSize[1] * lbSize[2];
        t.assign(N, 0); p.assign(N, 0); solid.assign(N, 0); exists.assign(N, 0);
        f.assign(size_t(N) * Q, 0.0); fs.assign(size_t(N) * Q, 0.0); d.assign(size_t(N) * Q, 0);
        n.assign(N, 0.0); u.assign(3 * N, 0.0); h.assign(3 * N, 0.0);
        mass.assign(N, 0.0); newMass.assign(N, 0.0); visc.assign(N, 0.0);
        // initializeLatticeBoundaries (LB.cpp:538-564)
        for (int64_t it = 0; it < N; ++it) {
            if (isWall(it)) continue;
            const unsigned x = getX(it), y = getY(it), z = getZ(it);
            if (x == 0) t[it] = boundary[0];
            else if (x == lbSize[0] - 1) t[it] = boundary[1];
            else if (y == 0) t[it] = boundary[2];
            else if (y == lbSize[1] - 1) t[it] = boundary[3];
            else if (z == 0) t[it] = boundary[4];
            else if (z == lbSize[2] - 1) t[it] = boundary[5];
        }
        // initializeTypes (LB.cpp:487-498): particle boundaries (activeNodes empty: only reset)
        initializeParticleBoundaries();
        // initializeWallBoundaries (LB.cpp:713-748)
        for (auto& w : walls) {
            const V3 wp = w.p / unit.Length;
            for (int64_t it = 0; it < N; ++it)
                if (getPosition(it).insidePlane(wp, w.n)) { solid[it] = uint16_t(w.index); t[it] = uint8_t(wallType); }
        }
        // initializeObjectBoundaries (LB.cpp:751-795)
        objectNodes.clear();
        for (auto& O : objects) {
            const V3 c = O.x0 / unit.Length; const double rr = O.r / unit.Length;
            for (int64_t it = 0; it < N; ++it)
                if (getPosition(it).insideSphere(c, rr)) { solid[it] = uint16_t(O.index); t[it] = 12; }
        }
        for (int64_t ip = 0; ip < N; ++ip) {
            if (t[ip] != 12) continue;
            createNode(ip);
            const Object& O = objects[solid[ip]];
            if (getPosition(ip).insideSphere(O.x0 / unit.Length, O.r / unit.Length)) {
                nodeInitialize(ip, initDensity, O.x1 / unit.Speed, 0.0, 0.0, V3());
                objectNodes.push_back(uint32_t(ip));
            }
        }
        if (lbRestart) {
            // restartInterface (LB.cpp:994-1009): the file's type at every fluid node
            for (int64_t it = 0; it < N; ++it)
                if (t[it] == 0) t[it] = restartTypes[size_t(it)];
            // restartVariables (LB.cpp:1098-1138): the next record at every active node, in index
            // order; node::restart (node.cpp:38-51) with initDensity added back
            size_t r = 0;
            const size_t nrec = restartRecords.size() / LBDEM_RESTART_REC;
            for (int64_t it = 0; it < N; ++it) {
                if (!isActive(it)) continue;
                if (r >= nrec) { restartShort = true; return; }
                const double* v = &restartRecords[r * LBDEM_RESTART_REC];
                ++r;
                createNode(it);
                n[it] = v[0] + initDensity;
                setU(it, V3(v[1], v[2], v[3]));
                mass[it] = v[4] + initDensity;
                visc[it] = v[5];
                for (int j = 0; j < Q; ++j) F(it)[j] = FS(it)[j] = v[6 + j];
            }
        } else {
            // initializeInterface (LB.cpp:974-992)
            if (freeSurface) {
                const double h0 = fluidHeight / unit.Length;
                for (int64_t it = 0; it < N; ++it)
                    if (getY(it) > h0 && t[it] == 0) t[it] = 2;
            }
            // initializeVariables (LB.cpp:1042-1096)
            for (int64_t it = 0; it < N; ++it) {
                if (t[it] == 0) {
                    createNode(it);
                    for (int j = 1; j < Q; ++j) {
                        const uint32_t l = d[size_t(it) * Q + j];
                        if (inRange(l) && t[l] == 2) { t[it] = 3; break; }
                    }
                }
                if (t[it] == 0) nodeInitialize(it, initDensity, initVelocity, initDensity, initDynVisc, lbF);
                else if (t[it] == 3) nodeInitialize(it, initDensity, initVelocity, 0.5 * initDensity, initDynVisc, lbF);
            }
        }
        // initializeWalls (LB.cpp:1140-1176)
        wallNodes.clear();
        for (int64_t it = 0; it < N; ++it) {
            if (isWall(it)) {
                createNode(it);
                // static and dynamic wall nodes alike: velocity zero (the dynamic one's solidVelocity
                // is never set, LB.cpp:1162-1169)
                if (t[it] >= 5 && t[it] <= 8) nodeInitialize(it, initDensity, V3(), 0.0, 0.0, V3());
                wallNodes.push_back(uint32_t(it));
            }
        }
        initializeLists();
        cleanLists();
        totalMass = 0.0;
        for (uint32_t idx : activeNodes) if (!p[idx]) totalMass += mass[idx];
    }

    // LB::initializeLists (LB.cpp:1011-1040)
    void initializeLists() {
        fluidNodes.clear(); particleNodes.clear(); interfaceNodes.clear();
        for (int64_t it = 0; it < N; ++it) {
            if (t[it] == 0) fluidNodes.push_back(uint32_t(it));
            else if (t[it] == 3) interfaceNodes.push_back(uint32_t(it));
            if (p[it]) particleNodes.push_back(uint32_t(it));
        }
    }
    static void sortDedupe(std::vector<uint32_t>& v) {
        std::sort(v.begin(), v.end());
        for (int ind = int(v.size()) - 1; ind > 0; --ind)
            if (v[ind] == v[ind - 1]) v.erase(v.begin() + ind);
    }
    // LB::cleanLists (LB.cpp:1181-1259)
    int cleanLists() {
        sortDedupe(fluidNodes); sortDedupe(interfaceNodes); sortDedupe(particleNodes); sortDedupe(objectNodes);
        activeNodes.clear();
        activeNodes.insert(activeNodes.end(), fluidNodes.begin(), fluidNodes.end());
        activeNodes.insert(activeNodes.end(), interfaceNodes.begin(), interfaceNodes.end());
        std::sort(activeNodes.begin(), activeNodes.end());
        for (int ind = int(activeNodes.size()) - 1; ind > 0; --ind) {
            const uint32_t i = activeNodes[ind], j = activeNodes[ind - 1];
            if (t[i] == 2 || t[i] == 12 || !exists[i] || i == j) activeNodes.erase(activeNodes.begin() + ind);
        }
        for (uint32_t a : activeNodes) if (t[a] == 2) return fail(LBDEM_E_STATE, "GAS ERROR: gas node in active list");
        return 0;
    }

    // ================= coupling (LB.cpp:432-467, 683-711, 797-943, 1875-1899, 1980-2100) =================
    // The reference tests every active node against every particle in particle order, so a node
    // inside several spheres keeps the highest index. Here each particle visits only the nodes of
    // its bounding box [floor(c - r), ceil(c + r)] that are on the active list, in the same
    // particle order: a node outside the box has one coordinate difference of at least r, so
    // its rounded squared distance is at least r*r and the strict test fails there anyway. The
    // flags are therefore the reference's (the fixtures compare them bit for bit).
    void initializeParticleBoundaries() {  // LB.cpp:683-711
        for (int64_t it = 0; it < N; ++it) { p[it] = 0; solid[it] = 0; }
        std::vector<uint8_t> onList(N, 0);
        for (uint32_t index : activeNodes) onList[index] = 1;
        const int64_t size[3] = {lbSize[0], lbSize[1], lbSize[2]};
        for (auto& P : particles) {
            const V3 c = P.x0 / unit.Length; const double rr = P.r / unit.Length;
            const unsigned idx = P.particleIndex;
            const double cc[3] = {c.x, c.y, c.z};
            int64_t lo[3], hi[3];
            bool empty = !(rr > 0.0);
            for (int a = 0; a < 3 && !empty; ++a) {
                if (!std::isfinite(cc[a] - rr) || !std::isfinite(cc[a] + rr)) { empty = true; break; }
                const double l = std::floor(cc[a] - rr), h = std::ceil(cc[a] + rr);
                if (h < 0.0 || l > double(size[a] - 1)) { empty = true; break; }
                lo[a] = l < 0.0 ? 0 : int64_t(l);
                hi[a] = h > double(size[a] - 1) ? size[a] - 1 : int64_t(h);
            }
            if (empty) continue;
            for (int64_t z = lo[2]; z <= hi[2]; ++z)
                for (int64_t y = lo[1]; y <= hi[1]; ++y)
                    for (int64_t x = lo[0]; x <= hi[0]; ++x) {
                        const int64_t index = x + y * size[0] + z * size[0] * size[1];
                        if (onList[index] && getPosition(index).insideSphere(c, rr)) {
                            p[index] = 1;
                            solid[index] = uint16_t(idx);
                        }
                    }
        }
    }
    void findNewActiveObject() {  // LB.cpp:846-943
        for (size_t ip = 0; ip < objectNodes.size(); ++ip) {
            const uint32_t index = objectNodes[ip];
            const V3 pos = getPosition(index);
            const Object& O = objects[0];
            bool newActive = true, surroundedGas = true;
            for (int j = 1; j < Q; ++j) {
                const uint32_t l = d[size_t(index) * Q + j];
                if (inRange(l) && t[l] == 0) { surroundedGas = false; break; }
            }
            if (pos.insideSphere(O.x0 / unit.Length, O.r / unit.Length)) newActive = false;
            if (newActive && surroundedGas) { t[index] = 2; deleteNode(index); newActive = false; }
            if (newActive) {
                createNode(index);  // created while still of object type: unwrapped links
                t[index] = 0;
                for (int j = 1; j < Q; ++j) {
                    const uint32_t l = d[size_t(index) * Q + j];
                    if (inRange(l) && t[l] == 2) { t[index] = 3; break; }
                }
                if (t[index] == 0) nodeInitialize(index, initDensity, initVelocity, initDensity, initDynVisc, lbF);
                else if (t[index] == 3) nodeInitialize(index, initDensity, V3(), 0.5 * initDensity, initDynVisc, lbF);
            }
        }
    }
    void updateObjectBoundaries() {  // LB.cpp:797-843
        findNewActiveObject();
        objectNodes.clear();
        for (auto& O : objects) {
            const V3 c = O.x0 / unit.Length; const double rr = O.r / unit.Length;
            for (int64_t it = 0; it < N; ++it)
                if (getPosition(it).insideSphere(c, rr)) { solid[it] = uint16_t(O.index); t[it] = 12; }
        }
        for (int64_t ip = 0; ip < N; ++ip) {
            if (t[ip] != 12) continue;
            createNode(ip);
            const Object& O = objects[solid[ip]];
            if (getPosition(ip).insideSphere(O.x0 / unit.Length, O.r / unit.Length)) {
                nodeInitialize(ip, initDensity, O.x1 / unit.Speed, 0.0, 0.0, V3());
                objectNodes.push_back(uint32_t(ip));
            }
        }
    }
    void findNewActive() {  // LB.cpp:1980-2029
        std::vector<uint32_t> popUp;
        for (size_t ip = 0; ip < particleNodes.size(); ++ip) {
            const uint32_t index = particleNodes[ip];
            const Particle& P = particles[solid[index]];
            if (getPosition(index).insideSphere(P.x0 / unit.Length, P.r / unit.Length)) break;
            popUp.push_back(index);
            p[index] = 0; solid[index] = 0;
        }
        for (uint32_t q : popUp)
            for (int ind = int(particleNodes.size()) - 1; ind >= 0; --ind)
                if (q == particleNodes[ind]) particleNodes.erase(particleNodes.begin() + ind);
    }
    // LB::findNewSolid (LB.cpp:2032-2100). The `break` at LB.cpp:2062 leaves the k-loop as
    // soon as a covered, unflagged neighbour is found, before the flagging block at
    // LB.cpp:2069-2094 can run; that block is therefore unreachable with newSolid == true and
    // the routine never changes a flag. Restated literally: scan, stop at the first hit.
    void findNewSolid() {
        for (size_t it = 0; it < particleNodes.size(); ++it) {
            const uint32_t index = particleNodes[it];
            const Particle& P = particles[solid[index]];
            const V3 c = P.x0 / unit.Length; const double rr = P.r / unit.Length;
            for (int k = 1; k < 7; ++k) {
                const uint32_t l = link(index, k);
                if (!inRange(l) || p[l]) continue;
                if (getPosition(l).insideSphere(c, rr)) break;
            }
        }
    }
    void couplingStep() {  // LB.cpp:432-467
        updateObjectBoundaries();
        if (newNeighborList) {
            initializeParticleBoundaries();  // updateIndices (LB.cpp:1875-1899)
            initializeLists();
            newNeighborList = false;
        }
        findNewActive();
        findNewSolid();
    }

    // ================= LB step (LB.cpp:400-418, 1261-1544, 1901-1978) =================
    void computeHydroForces() {  // LB.cpp:1901-1978
        for (auto& P : particles) { P.FHydro = V3(); P.MHydro = V3(); P.fluidVolume = 0.0; P.SHydro.reset(); }
        for (uint32_t index : activeNodes) {
            setH(index, V3());
            if (!p[index]) continue;
            Particle& P = particles[solid[index]];
            const V3 radius = getPosition(index) - P.x0 / unit.Length;
            const V3 localVel = P.x1 / unit.Speed + (P.w1.cross(radius)) / unit.AngVel;
            const V3 diffVel = (mass[index] / n[index]) * (U(index) - localVel);
            V3 hh = H(index); hh += -1.0 * diffVel; setH(index, hh);
            P.fluidVolume += mass[index];
            P.FHydro += 1.0 * diffVel;
            P.MHydro += 1.0 * radius.cross(diffVel);
            const double rr[3] = {radius.x, radius.y, radius.z}, dd[3] = {diffVel.x, diffVel.y, diffVel.z};
            for (int a = 0; a < 3; ++a)  // LB.cpp:1951-1959
                for (int b = 0; b < 3; ++b) P.SHydro.m[3 * a + b] += 1.0 * rr[a] * unit.Length * dd[b] * unit.Force;
        }
        for (auto& P : particles) { P.FHydro *= unit.Force; P.MHydro *= unit.Torque; P.fluidVolume *= unit.Volume; }
    }
    // node.cpp:113-168: shear rate from the non-equilibrium part relaxed with the node's previous
    // viscosity, apparent viscosity of the rheology model plus the Smagorinsky term, clamped to the
    // viscosities of tau in [minTau, maxTau] (lattice.h:26-27) with std::min / std::max
    void apparentViscosity(uint32_t index, const double* fi, const double* feq, const V3& uu) {
        const double minVisc = (0.500001 - 0.5) / 3.0, maxVisc = (1.8 - 0.5) / 3.0;
        const double tau = 0.5 + 3.0 * visc[index];
        double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int j = 0; j < Q; ++j) {
            const V3 vj = vdir(j);
            const double c[3] = {vj.x, vj.y, vj.z};
            const double d = fi[j] - feq[j];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) m[a][b] += d * (c[a] * c[b]);
        }
        const double sc = 1.5 / (tau * n[index]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) m[a][b] *= sc;
        const double mag = std::sqrt(0.5 * (m[0][0] * m[0][0] + m[1][1] * m[1][1] + m[2][2] * m[2][2] +
                                            2.0 * (m[0][1] * m[1][0] + m[2][0] * m[0][2] + m[1][2] * m[2][1])));
        const double shearRate = 2.0 * mag;
        double nuApp = 0.0;
        switch (rheologyModel) {
            case LBDEM_RHEO_BINGHAM: nuApp = plasticVisc + yieldStress / shearRate; break;
            case LBDEM_RHEO_FRICTIONAL: {
                const double pressure = std::max(0.0, 0.33333333 * (n[index] - 1.0));
                nuApp = initDynVisc + frictionCoefFluid * pressure / shearRate;
                break;
            }
            case LBDEM_RHEO_VOELLMY: {
                const double pressure = std::max(0.0, 0.33333333 * (n[index] - 1.0));
                nuApp = (frictionCoefFluid * pressure + uu.norm2() / voellmyCoefficient) / shearRate;
                break;
            }
            default: nuApp = initDynVisc; break;
        }
        if (turbulenceOn) nuApp += turbConst * shearRate;
        visc[index] = std::max(minVisc, std::min(maxVisc, nuApp));
    }
    void collision() {  // LB.cpp:1301-1332 with node.cpp:95-98, 181-207
        if (!forceField) lbF.reset();
        const int64_t na = int64_t(activeNodes.size());
#pragma omp parallel for schedule(static)
        for (int64_t it = 0; it < na; ++it) {
            const uint32_t index = activeNodes[it];
            double* fi = F(index);
            const V3 totalForce = lbF + H(index);
            V3 uu = U(index);
            uu += 0.5 * totalForce / n[index];
            setU(index, uu);
            double feq[Q];
            equilibrium(n[index], uu, feq);
            if (rheologyModel != LBDEM_RHEO_NEWTONIAN || turbulenceOn) apparentViscosity(index, fi, feq, uu);
            const double omega = 1.0 / (0.5 + 3.0 * visc[index]);
            for (int j = 0; j < Q; j++) fi[j] += omega * (feq[j] - fi[j]);
            const double omegaf = 1.0 - 1.0 / (1.0 + 6.0 * visc[index]);
            for (int j = 0; j < Q; j++) {
                const V3 vj = vdir(j);
                const double vu = uu.dot(vj);
                const V3 vmu = vj - uu;
                const V3 forcePart = 9.0 * vu * vj + 3.0 * vmu;
                fi[j] += omegaf * coeff(j) * forcePart.dot(totalForce);
            }
        }
    }
    int streaming() {  // LB.cpp:1334-1544 (types 0/3 active, 2 gas, 7 static wall, 12 object)
        double staticPres[Q];
        for (int j = 0; j < Q; ++j) staticPres[j] = initDensity * coeff(j);
        for (auto& w : walls) w.FHydro.reset();
        for (auto& o : objects) { o.FHydro.reset(); o.MHydro.reset(); }
        const int64_t na = int64_t(activeNodes.size());
#pragma omp parallel for schedule(static)
        for (int64_t it = 0; it < na; ++it) {
            const uint32_t i = activeNodes[it];
            for (int j = 0; j < Q; ++j) FS(i)[j] = F(i)[j];
        }
        int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
        for (int64_t in = 0; in < na; ++in) {
            const uint32_t it = activeNodes[in];
            const double* fsi = FS(it);
            double* fi = F(it);
            for (int j = 1; j < Q; ++j) {
                const uint32_t l = d[size_t(it) * Q + j];
                const uint8_t tl = t[l];
                if (tl == 0 || tl == 3) fi[OPP[j]] = FS(l)[OPP[j]];
                else if (tl == 2) {
                    const V3 uu = U(it);
                    const double usq = uu.norm2(), vuj = uu.dot(vdir(j));
                    fi[OPP[j]] = -fsi[j] + coeff(j) * initDensity * (2.0 + 9.0 * (vuj * vuj) - 3.0 * usq);
                } else if (tl == 7) fi[OPP[j]] = fsi[j];
                else if (tl == 5) {  // static slip wall (LB.cpp:1454-1482)
                    const double S1 = slipCoefficient, S2 = (1.0 - slipCoefficient);
                    if (j > 6) {
                        const uint32_t c1 = d[size_t(it) * Q + SLIP1_CHECK[j]], c2 = d[size_t(it) * Q + SLIP2_CHECK[j]];
                        const bool a1 = isActive(c1), a2 = isActive(c2);
                        if (a1 && !a2) fi[OPP[j]] = S1 * FS(c1)[SLIP1[j]] + S2 * fsi[j];
                        else if (!a1 && a2) fi[OPP[j]] = S1 * FS(c2)[SLIP2[j]] + S2 * fsi[j];
                        else fi[OPP[j]] = fsi[j];
                    } else {
                        fi[OPP[j]] = fsi[j];
                    }
                } else if (tl == 6) {  // moving slip wall (LB.cpp:1483-1523): the wall node's velocity
                    const double S1 = slipCoefficient, S2 = (1.0 - slipCoefficient);
                    const double BBi = 6.0 * n[it] * coeff(j) * U(l).dot(vdir(j));
                    if (j > 6) {
                        const uint32_t c1 = d[size_t(it) * Q + SLIP1_CHECK[j]], c2 = d[size_t(it) * Q + SLIP2_CHECK[j]];
                        const bool a1 = isActive(c1), a2 = isActive(c2);
                        if (a1 && !a2) fi[OPP[j]] = S1 * FS(c1)[SLIP1[j]] + S2 * (fsi[j] - BBi);
                        else if (!a1 && a2) fi[OPP[j]] = S1 * FS(c2)[SLIP2[j]] + S2 * (fsi[j] - BBi);
                        else fi[OPP[j]] = fsi[j] - BBi;
                    } else {
                        fi[OPP[j]] = fsi[j] - BBi;
                    }
                } else if (tl == 8 || tl == 12) {  // moving wall or object (LB.cpp:1403-1421, 1435-1453)
                    const V3 vel = U(l);
                    const double BBi = 6.0 * n[it] * coeff(j) * vel.dot(vdir(j));
                    fi[OPP[j]] = fsi[j] - BBi;
                } else bad++;
            }
        }
        if (bad) return fail(LBDEM_E_TYPE, "TYPE ERROR: streaming met an unsupported neighbour type");
        // momentum exchange and extra mass accumulated in ascending (node, direction) order,
        // the order the reference's `omp ordered` sections impose (LB.cpp:1414-1453)
        double extraMass = 0.0;
        for (int64_t in = 0; in < na; ++in) {
            const uint32_t it = activeNodes[in];
            const double* fsi = FS(it);
            for (int j = 1; j < Q; ++j) {
                const uint32_t l = d[size_t(it) * Q + j];
                if (t[l] == 7) {
                    const V3 BBforce = (2.0 * (fsi[j] - staticPres[j]) - 0.0) * vdir(j);
                    const unsigned si = solid[l];
                    walls[si < walls.size() ? si : 0].FHydro += BBforce;
                } else if (t[l] == 6) {  // moving slip wall: extra mass only (LB.cpp:1483-1523)
                    const double S2 = (1.0 - slipCoefficient);
                    const double BBi = 6.0 * n[it] * coeff(j) * U(l).dot(vdir(j));
                    bool slip = false;
                    if (j > 6) {
                        const bool a1 = isActive(d[size_t(it) * Q + SLIP1_CHECK[j]]);
                        const bool a2 = isActive(d[size_t(it) * Q + SLIP2_CHECK[j]]);
                        slip = a1 != a2;
                    }
                    if (slip) extraMass += S2 * mass[it] * BBi;
                    else extraMass += mass[it] * BBi;
                } else if (t[l] == 8) {  // moving wall (LB.cpp:1403-1421): the wall node's velocity
                    const unsigned si = solid[l];
                    const V3 vel = U(l);
                    const double BBi = 6.0 * n[it] * coeff(j) * vel.dot(vdir(j));
                    const V3 BBforce = (2.0 * (fsi[j] - staticPres[j]) - BBi) * vdir(j);
                    walls[si < walls.size() ? si : 0].FHydro += BBforce;
                    extraMass += BBi * mass[it];
                } else if (t[l] == 12) {
                    const unsigned si = solid[l];
                    const V3 vel = U(l);
                    const double BBi = 6.0 * n[it] * coeff(j) * vel.dot(vdir(j));
                    const V3 BBforce = (2.0 * (fsi[j] - staticPres[j]) - BBi) * vdir(j);
                    if (si < objects.size()) {  // pinned: see header
                        const V3 rad = getPosition(l) - objects[si].x0 / unit.Length;
                        objects[si].FHydro += BBforce;
                        objects[si].MHydro += rad.cross(BBforce);
                    }
                    extraMass += BBi * mass[it];
                }
            }
        }
        lastExtraMass = extraMass;
        redistributeMass(extraMass);
        for (auto& w : walls) w.FHydro *= unit.Force;
        for (auto& o : objects) { o.FHydro *= unit.Force; o.MHydro *= unit.Torque; }
        return 0;
    }
    int lbStep() {  // LB.cpp:400-418
        initializeLists();
        int rc = cleanLists();
        if (rc) return rc;
        const int64_t na = int64_t(activeNodes.size());
#pragma omp parallel for schedule(static)
        for (int64_t it = 0; it < na; ++it) {  // reconstruct (node.cpp:75-93)
            const uint32_t i = activeNodes[it];
            const double* fi = F(i);
            double nn = 0.0;
            for (int j = 0; j < Q; ++j) nn += fi[j];
            V3 uu;
            for (int j = 0; j < Q; ++j) uu += fi[j] * vdir(j);
            uu /= nn;
            n[i] = nn; setU(i, uu);
        }
        computeHydroForces();
        collision();
        return streaming();
    }

    // ================= free surface (LB.cpp:422-430, 1548-1852) =================
    void redistributeMass(double massSurplus) {  // LB.cpp:1844-1852
        const double addMass = massSurplus / double(interfaceNodes.size());
        for (uint32_t i : interfaceNodes) mass[i] += addMass;
    }
    void updateMass() {  // LB.cpp:1548-1639
        for (uint32_t i : interfaceNodes) newMass[i] = mass[i];
        for (uint32_t index : interfaceNodes) {
            double deltaMass = 0.0;
            const double* fi = F(index);
            const double* fsi = FS(index);
            for (int j = 1; j < Q; ++j) {
                const uint32_t l = d[size_t(index) * Q + j];
                const double ms = fi[OPP[j]] - fsi[j];
                if (t[l] == 3) deltaMass += 0.5 * (mass[l] + mass[index]) * ms;
                else if (t[l] == 0) deltaMass += 1.0 * ms;
                else if (t[l] == 8 || t[l] == 12) deltaMass += 1.0 * mass[index] * ms;
                else if (t[l] == 6) {  // moving slip wall (LB.cpp:1595-1625)
                    double averageMass = 0.0;
                    if (j > 6) {
                        const bool a1 = isActive(d[size_t(index) * Q + SLIP1_CHECK[j]]);
                        const bool a2 = isActive(d[size_t(index) * Q + SLIP2_CHECK[j]]);
                        if (a1 != a2) averageMass += 1.0 * (1.0 - slipCoefficient) * mass[index];
                        else averageMass += 1.0 * mass[index];
                    } else {
                        averageMass = 1.0 * mass[index];
                    }
                    deltaMass += averageMass * ms;
                }
            }
            newMass[index] += deltaMass;
        }
        for (uint32_t i : fluidNodes) mass[i] = n[i];
        for (uint32_t i : interfaceNodes) mass[i] = newMass[i];
    }
    void updateInterface() {  // LB.cpp:1641-1667
        double massSurplus = 0.0;
        std::vector<uint32_t> filled, emptied, newIf;
        // findInterfaceMutants (LB.cpp:1669-1698)
        for (int ind = int(interfaceNodes.size()) - 1; ind >= 0; --ind) {
            const uint32_t i = interfaceNodes[ind];
            if (mass[i] > n[i]) {
                filled.push_back(i); t[i] = 0; fluidNodes.push_back(i);
                interfaceNodes.erase(interfaceNodes.begin() + ind);
            } else if (mass[i] < 0.0) {
                emptied.push_back(i); t[i] = 2;
                interfaceNodes.erase(interfaceNodes.begin() + ind);
            }
        }
        // smoothenInterface (LB.cpp:1700-1749)
        for (uint32_t index : filled) {
            for (int j = 1; j < Q; ++j) {
                const uint32_t l = d[size_t(index) * Q + j];
                if (t[l] == 2) {
                    t[l] = 3;
                    newIf.push_back(l);
                    createNode(l);
                    nodeInitialize(l, initDensity, U(index), 0.01 * initDensity, visc[index], H(index) + lbF);
                    massSurplus -= 0.01 * initDensity;
                }
            }
        }
        for (uint32_t index : emptied) {
            for (int j = 0; j < Q; ++j) {
                const uint32_t l = d[size_t(index) * Q + j];
                if (t[l] == 0) {
                    t[l] = 3;
                    newIf.push_back(l);
                    mass[l] = 0.99 * n[l];
                    massSurplus += 0.01 * n[l];
                    for (int ind = int(fluidNodes.size()) - 1; ind >= 0; --ind)
                        if (fluidNodes[ind] == l) fluidNodes.erase(fluidNodes.begin() + ind);
                }
            }
        }
        // updateMutants (LB.cpp:1751-1790)
        for (uint32_t e : emptied) {
            if (std::find(newIf.begin(), newIf.end(), e) == newIf.end()) { massSurplus += mass[e]; deleteNode(e); }
        }
        for (uint32_t fl : filled) {
            if (std::find(newIf.begin(), newIf.end(), fl) == newIf.end()) { massSurplus += mass[fl] - n[fl]; mass[fl] = n[fl]; }
        }
        for (uint32_t q : newIf) interfaceNodes.push_back(q);
        // removeIsolated (LB.cpp:1792-1842)
        for (int ind = int(interfaceNodes.size()) - 1; ind >= 0; --ind) {
            const uint32_t i = interfaceNodes[ind];
            bool surroundedFluid = true;
            for (int j = 1; j < Q; ++j) if (t[d[size_t(i) * Q + j]] == 2) { surroundedFluid = false; break; }
            if (surroundedFluid) {
                massSurplus += mass[i] - n[i];
                mass[i] = n[i];
                t[i] = 0;
                interfaceNodes.erase(interfaceNodes.begin() + ind);
                fluidNodes.push_back(i);
            }
        }
        for (int ind = int(interfaceNodes.size()) - 1; ind >= 0; --ind) {
            const uint32_t i = interfaceNodes[ind];
            bool surroundedGas = true;
            for (int j = 1; j < Q; ++j) if (t[d[size_t(i) * Q + j]] == 0) { surroundedGas = false; break; }
            if (surroundedGas) {
                massSurplus += mass[i];
                interfaceNodes.erase(interfaceNodes.begin() + ind);
                t[i] = 2;
                deleteNode(i);
            }
        }
        lastSurplus = massSurplus;
        redistributeMass(massSurplus);
    }
    int fsStep() {  // LB.cpp:422-430
        updateMass();
        updateInterface();
        return cleanLists();
    }

    // ================= DEM (DEM.cpp:359-2836) =================
    int neighborList() {  // DEM.cpp:473-652
        nlBuilds++;
        const int num = int(particles.size());
        for (auto& P : particles) {
            std::memcpy(P._list, P.list, sizeof(P.list));
            std::memcpy(P._xix, P.xix, sizeof(P.xix));
            std::memcpy(P._xiy, P.xiy, sizeof(P.xiy));
            std::memcpy(P._xiz, P.xiz, sizeof(P.xiz));
        }
        double r_max = 0.0;
        for (auto& P : particles) {
            r_max = dmax(r_max, P.r);
            P.x0_lst = P.x0;
            P.list[0] = 0; P.xix[0] = P.xiy[0] = P.xiz[0] = 0.0;
        }
        double cut = 2.0 * r_max + 2.0;
        const int nlx = int(demSize[0] / cut), nly = int(demSize[1] / cut), nlz = int(demSize[2] / cut);
        bool overflow = false;
        auto add = [&](int n1, int n2) {
            Particle& A = particles[n1]; Particle& B = particles[n2];
            if (A.list[0] + 1 >= MAXL || B.list[0] + 1 >= MAXL) { overflow = true; return false; }
            A.list[0]++; B.list[0]++;
            A.xix[0] += 1.0; A.xiy[0] += 1.0; A.xiz[0] += 1.0;
            B.xix[0] += 1.0; B.xiy[0] += 1.0; B.xiz[0] += 1.0;
            A.list[A.list[0]] = n2; B.list[B.list[0]] = n1;
            return true;
        };
        auto within = [&](int n1, int n2) {
            const Particle& A = particles[n1]; const Particle& B = particles[n2];
            const double x12 = n_image(A.x0.x - B.x0.x, demSize[0]);
            const double y12 = n_image(A.x0.y - B.x0.y, demSize[1]);
            const double z12 = n_image(A.x0.z - B.x0.z, demSize[2]);
            const double r12 = x12 * x12 + y12 * y12 + z12 * z12;
            const double sig = A.r + B.r;
            const double c = sig + 2.0;
            return r12 <= c * c;
        };
        if ((nlx < nly ? nlx : nly) < 3) {
            for (int n1 = 1; n1 < num; n1++)
                for (int n2 = 0; n2 < n1; n2++) {
                    if (!within(n1, n2)) continue;
                    if (!add(n1, n2)) continue;
                    Particle& A = particles[n1]; Particle& B = particles[n2];
                    A.xix[A.list[0]] = B.xix[B.list[0]]; B.xix[B.list[0]] = A.xix[A.list[0]];
                    A.xiy[A.list[0]] = B.xiy[B.list[0]]; B.xiy[B.list[0]] = A.xiy[A.list[0]];
                    A.xiz[A.list[0]] = B.xiz[B.list[0]]; B.xiz[B.list[0]] = A.xiz[A.list[0]];
                }
        } else {
            // nlz may be 0, 1 or 2 here (only nlx and nly are tested, DEM.cpp:515). With 1 or 2
            // z-cells the wrapped walk below visits a z-cell more than once and lists a pair
            // once per visit, as the reference does; with 0 its loops never run and every list
            // stays empty (the reference then files every particle under cell 0, which nothing
            // reads).
            const int ncell = nlx * nly * nlz;
            if (ncell == 0) return 0;  // every list was emptied above
            std::vector<std::vector<int>> cell(ncell);
            for (auto& P : particles)
                if (!std::isfinite(P.x0.x) || !std::isfinite(P.x0.y) || !std::isfinite(P.x0.z)) {
                    err = "particle position is not finite";
                    return LBDEM_E_STATE;
                }
            for (int n1 = 0; n1 < num; n1++) {
                Particle& P = particles[n1];
                const int lx = int(dbox(P.x0.x / demSize[0], 1) * nlx);
                const int ly = int(dbox(P.x0.y / demSize[1], 1) * nly);
                const int lz = int(dbox(P.x0.z / demSize[2], 1) * nlz);
                cell[lx * nly * nlz + ly * nlz + lz].push_back(n1);
                P.list[0] = 0; P.xix[0] = P.xiy[0] = P.xiz[0] = 0.0;
            }
            for (int lx1 = 0; lx1 < nlx; lx1++)
                for (int ly1 = 0; ly1 < nly; ly1++)
                    for (int lz1 = 0; lz1 < nlz; lz1++)
                        for (int lx2 = lx1 - 1; lx2 <= lx1 + 1; lx2++)
                            for (int ly2 = ly1 - 1; ly2 <= ly1 + 1; ly2++)
                                for (int lz2 = lz1 - 1; lz2 <= lz1 + 1; lz2++) {
                                    const int l1 = lx1 * nly * nlz + ly1 * nlz + lz1;
                                    const int l2 = ((lx2 + nlx) % nlx) * nly * nlz + ((ly2 + nly) % nly) * nlz + (lz2 + nlz) % nlz;
                                    for (int a : cell[l1])
                                        for (int b : cell[l2])
                                            if (b < a && within(a, b)) add(a, b);
                                }
        }
        if (overflow) { err = "neighbour list capacity (63) exceeded"; }
        for (auto& P : particles)
            for (int nn = 1; nn <= P.list[0]; nn++) {
                const int k = P.list[nn];
                for (int l = 1; l <= P._list[0]; l++)
                    if (P._list[l] == k) { P.xix[nn] = P._xix[l]; P.xiy[nn] = P._xiy[l]; P.xiz[nn] = P._xiz[l]; }
            }
        return overflow ? LBDEM_E_CAPACITY : 0;
    }

    // Jeffrey-Onishi / Nguyen-Ladd lubrication coefficients shared by the particle-particle
    // and object-particle routines (DEM.cpp:1181-1257, 2349-2416)
    // S: xx yy zz yz zx xy; normal: within the normal range (DEM.cpp:2360-2363 counts the pair)
    struct LubOut { V3 F, T1, T2; double S1[6] = {0, 0, 0, 0, 0, 0}, S2[6] = {0, 0, 0, 0, 0, 0}; bool normal = false; };
    LubOut pairLub(const V3& xa, const V3& xb, double ra, double rb, const V3& va, const V3& vb, const V3& wa, const V3& wb,
                   double beta1_num, double beta1_den, bool& active) const {
        LubOut o;
        const double lub_N = prm.lubNormal, lub_T = prm.lubTangential, lub_R = prm.lubRotation;
        const double cut = prm.lubCutoff;
        const double rad = 0.5 * (ra + rb);
        double x12 = xa.x - xb.x, y12 = xa.y - xb.y, z12 = xa.z - xb.z;
        const double r12 = std::sqrt(x12 * x12 + y12 * y12 + z12 * z12);
        const double umx = va.x - vb.x, umy = va.y - vb.y, umz = va.z - vb.z;
        const double w1x = wa.x, w1y = wa.y, w1z = wa.z, w2x = wb.x, w2y = wb.y, w2z = wb.z;
        const double dr = r12 - 2.0 * rad;
        const double h = dmax(dr, cut);
        active = dr < dmax(cut, dmax(lub_N, dmax(lub_T, lub_R)));
        if (!active) return o;
        double Xa, Ya, Yb1, Yb2, Yc11, Yc12, Yc21, Yc22, h_c, beta;
        double Xg1, Xg2, Yg1, Yg2, Yh11, Yh12, Yh21, Yh22;  // stresslet coefficients (DEM.cpp:2362-2406)
        const double nu = prm.kinVisc;
        const double alpha = M_PI * nu * rad * deltat;
        const double beta1 = beta1_num / beta1_den;
        const double beta2 = beta1 * beta1, beta3 = beta1 * beta2;
        const double beta4 = std::pow((1 + beta1), 4.0);
        const double beta5 = std::pow((1 + beta1), 5.0);
        if (dr < dmax(cut, lub_N)) {
            o.normal = true;
            Xa = 4.0 * beta2 / beta4;
            Xg1 = 12.0 * beta2 / beta5;
            Xg2 = -12.0 * beta3 / beta5;
            h_c = dmax(dr, lub_N);
            Xa *= rad / h - rad / h_c;
            Xg1 *= rad / h - rad / h_c;
            Xg2 *= rad / h - rad / h_c;
            Xa *= 6.0 * alpha;
            Xg1 *= 4.0 * alpha * rad;
            Xg2 *= 4.0 * alpha * rad;
        } else Xa = Xg1 = Xg2 = 0.0;
        if (dr < dmax(cut, lub_T)) {
            Ya = 8.0 * beta1 * (2.0 + beta1 + 2.0 * beta2) / (15.0 * beta4);
            Yb1 = -4.0 * beta1 * (4.0 + beta1) / (5.0 * beta4);
            Yb2 = 4.0 * beta1 * (4.0 * beta2 + beta1) / (5.0 * beta4);
            Yg1 = 2.0 * beta1 * (4.0 - beta1 + 7.0 * beta2) / (5.0 * beta5);
            Yg2 = -2.0 * beta2 * (4.0 * beta2 - beta1 + 7.0) / (5.0 * beta5);
            h_c = dmax(dr, lub_T);
            beta = alpha * std::log(h_c / h);
            Ya *= 6.0 * beta;
            Yb1 *= 4.0 * beta * rad;
            Yb2 *= 4.0 * beta * rad;
            Yg1 *= 4.0 * beta * rad;
            Yg2 *= 4.0 * beta * rad;
        } else Ya = Yb1 = Yb2 = Yg1 = Yg2 = 0.0;
        if (dr < dmax(cut, lub_R)) {
            Yc11 = 16.0 * beta1 / (5.0 * beta4);
            Yc12 = 4.0 * beta2 / (5.0 * beta4);
            Yc21 = 4.0 * beta2 / (5.0 * beta4);
            Yc22 = 16.0 * beta3 / (5.0 * beta4);
            Yh11 = 4.0 * beta1 * (2.0 - beta1) / (5.0 * beta5);
            Yh12 = 2.0 * beta2 * (1.0 + 7.0 * beta1) / (5.0 * beta5);
            Yh21 = 2.0 * beta2 * (beta1 + 7.0) / (5.0 * beta5);
            Yh22 = 4.0 * beta2 * (2.0 * beta2 - beta1) / (5.0 * beta5);
            h_c = dmax(dr, lub_R);
            beta = alpha * std::log(h_c / h);
            Yc11 *= 8.0 * beta * rad * rad;
            Yc12 *= 8.0 * beta * rad * rad;
            Yc21 *= 8.0 * beta * rad * rad;
            Yc22 *= 8.0 * beta * rad * rad;
            Yh11 *= 8.0 * beta * rad * rad;
            Yh21 *= 8.0 * beta * rad * rad;
            Yh12 *= 8.0 * beta * rad * rad;
            Yh22 *= 8.0 * beta * rad * rad;
        } else Yc11 = Yc12 = Yc21 = Yc22 = Yh11 = Yh12 = Yh21 = Yh22 = 0.0;
        x12 /= r12; y12 /= r12; z12 /= r12;
        const double um_r = umx * x12 + umy * y12 + umz * z12;
        const double w1_r = w1x * x12 + w1y * y12 + w1z * z12;
        const double w2_r = w2x * x12 + w2y * y12 + w2z * z12;
        const double umxrx = umy * z12 - umz * y12, umxry = umz * x12 - umx * z12, umxrz = umx * y12 - umy * x12;
        const double w1xrx = w1y * z12 - w1z * y12, w1xry = w1z * x12 - w1x * z12, w1xrz = w1x * y12 - w1y * x12;
        const double w2xrx = w2y * z12 - w2z * y12, w2xry = w2z * x12 - w2x * z12, w2xrz = w2x * y12 - w2y * x12;
        double Fx = -Xa * um_r * x12 - Ya * (umx - um_r * x12);
        double Fy = -Xa * um_r * y12 - Ya * (umy - um_r * y12);
        double Fz = -Xa * um_r * z12 - Ya * (umz - um_r * z12);
        Fx += -Yb1 * w1xrx + Yb2 * w2xrx;
        Fy += -Yb1 * w1xry + Yb2 * w2xry;
        Fz += -Yb1 * w1xrz + Yb2 * w2xrz;
        double T1x = Yb1 * umxrx, T1y = Yb1 * umxry, T1z = Yb1 * umxrz;
        double T2x = -Yb2 * umxrx, T2y = -Yb2 * umxry, T2z = -Yb2 * umxrz;
        T1x += -Yc11 * (w1x - w1_r * x12) - Yc12 * (w2x - w2_r * x12);
        T1y += -Yc11 * (w1y - w1_r * y12) - Yc12 * (w2y - w2_r * y12);
        T1z += -Yc11 * (w1z - w1_r * z12) - Yc12 * (w2z - w2_r * z12);
        T2x += -Yc21 * (w1x - w1_r * x12) - Yc22 * (w2x - w2_r * x12);
        T2y += -Yc21 * (w1y - w1_r * y12) - Yc22 * (w2y - w2_r * y12);
        T2z += -Yc21 * (w1z - w1_r * z12) - Yc22 * (w2z - w2_r * z12);
        o.F = V3(Fx, Fy, Fz); o.T1 = V3(T1x, T1y, T1z); o.T2 = V3(T2x, T2y, T2z);
        // stresslets of the two spheres (DEM.cpp:2454-2477), as xx yy zz yz zx xy
        double* S1 = o.S1;
        double* S2 = o.S2;
        S1[0] = Xg1 * um_r * x12 * x12 + Yg1 * (x12 * umx + umx * x12);
        S1[1] = Xg1 * um_r * y12 * y12 + Yg1 * (y12 * umy + umy * y12);
        S1[2] = Xg1 * um_r * z12 * z12 + Yg1 * (z12 * umz + umz * z12);
        S1[3] = Xg1 * um_r * y12 * z12 + Yg1 * (y12 * umz + umy * z12);
        S1[4] = Xg1 * um_r * z12 * x12 + Yg1 * (z12 * umx + umz * x12);
        S1[5] = Xg1 * um_r * x12 * y12 + Yg1 * (x12 * umy + umx * y12);
        S2[0] = -Xg2 * um_r * x12 * x12 - Yg2 * (x12 * umx + umx * x12);
        S2[1] = -Xg2 * um_r * y12 * y12 - Yg2 * (y12 * umy + umy * y12);
        S2[2] = -Xg2 * um_r * z12 * z12 - Yg2 * (z12 * umz + umz * z12);
        S2[3] = -Xg2 * um_r * y12 * z12 - Yg2 * (y12 * umz + umy * z12);
        S2[4] = -Xg2 * um_r * z12 * x12 - Yg2 * (z12 * umx + umz * x12);
        S2[5] = -Xg2 * um_r * x12 * y12 - Yg2 * (x12 * umy + umx * y12);
        S1[0] += -Yh11 * (x12 * w1xrx + w1xrx * x12) - Yh12 * (x12 * w2xrx + w2xrx * x12);
        S1[1] += -Yh11 * (y12 * w1xry + w1xry * y12) - Yh12 * (y12 * w2xry + w2xry * y12);
        S1[2] += -Yh11 * (z12 * w1xrz + w1xrz * z12) - Yh12 * (z12 * w2xrz + w2xrz * z12);
        S1[3] += -Yh11 * (y12 * w1xrz + w1xry * z12) - Yh12 * (y12 * w2xrz + w2xry * z12);
        S1[4] += -Yh11 * (z12 * w1xrx + w1xrz * x12) - Yh12 * (z12 * w2xrx + w2xrz * x12);
        S1[5] += -Yh11 * (x12 * w1xry + w1xrx * y12) - Yh12 * (x12 * w2xry + w2xrx * y12);
        S2[0] += -Yh21 * (x12 * w1xrx + w1xrx * x12) - Yh22 * (x12 * w2xrx + w2xrx * x12);
        S2[1] += -Yh21 * (y12 * w1xry + w1xry * y12) - Yh22 * (y12 * w2xry + w2xry * y12);
        S2[2] += -Yh21 * (z12 * w1xrz + w1xrz * z12) - Yh22 * (z12 * w2xrz + w2xrz * z12);
        S2[3] += -Yh21 * (y12 * w1xrz + w1xry * z12) - Yh22 * (y12 * w2xrz + w2xry * z12);
        S2[4] += -Yh21 * (z12 * w1xrx + w1xrz * x12) - Yh22 * (z12 * w2xrx + w2xrz * x12);
        S2[5] += -Yh21 * (x12 * w1xry + w1xrx * y12) - Yh22 * (x12 * w2xry + w2xrx * y12);
        return o;
    }
    void particleLubrication() {  // DEM.cpp:2302-2521
        for (int n1 = 0; n1 < int(particles.size()); n1++) {
            Particle& A = particles[n1];
            for (int nn = 1; nn <= A.list[0]; nn++) {
                const int n2 = A.list[nn];
                if (n2 >= n1) continue;
                Particle& B = particles[n2];
                bool act;
                LubOut o = pairLub(A.x0, B.x0, A.r, B.r, A.x1, B.x1, A.w1, B.w1, B.r, A.r, act);
                if (!act) continue;
                A.FLub += o.F; B.FLub -= o.F; A.MLub += o.T1; B.MLub += o.T2;
                A.SLub.addSym(o.S1); B.SLub.addSym(o.S2);
                if (o.normal) { A.lub_coord += 1.0; B.lub_coord += 1.0; }
            }
        }
    }
    void wallLubrication() {  // DEM.cpp:2525-2672
        const double lub_N = prm.lubNormal, lub_T = prm.lubTangential, lub_R = prm.lubRotation, cut = prm.lubCutoff;
        for (auto& P : particles)
            for (auto& W : walls) {
                const double rad = P.r;
                double x12 = (P.x0.x - W.p.x) * W.n.x * W.n.x;
                double y12 = (P.x0.y - W.p.y) * W.n.y * W.n.y;
                double z12 = (P.x0.z - W.p.z) * W.n.z * W.n.z;
                const double r12 = std::sqrt(x12 * x12 + y12 * y12 + z12 * z12);
                const double umx = P.x1.x, umy = P.x1.y, umz = P.x1.z;
                const double w1x = P.w1.x, w1y = P.w1.y, w1z = P.w1.z;
                const double dr = r12 - rad;
                const double h = dmax(dr, cut);
                if (!(dr < dmax(cut, dmax(lub_N, dmax(lub_T, lub_R))))) continue;
                const double alpha = M_PI * prm.kinVisc * rad * deltat;
                double Xa, Ya, Yb1, Yc11, h_c, beta, Xg1, Yg1, Yh11;
                if (dr < dmax(cut, lub_N)) {
                    Xa = 1.0; Xg1 = 3.0 / 2.0; h_c = dmax(dr, lub_N);
                    Xa *= rad / h - rad / h_c; Xg1 *= rad / h - rad / h_c;
                    Xa *= 6.0 * alpha; Xg1 *= 4.0 * alpha * rad;
                } else Xa = Xg1 = 0.0;
                if (dr < dmax(cut, lub_T)) {
                    Ya = 8.0 / 15.0; Yb1 = -1.0 / 5.0; Yg1 = 7.0 / 10.0; h_c = dmax(dr, lub_T); beta = alpha * std::log(h_c / h);
                    Ya *= 6.0 * beta; Yb1 *= 4.0 * beta * rad; Yg1 *= 4.0 * beta * rad;
                } else Ya = Yb1 = Yg1 = 0.0;
                if (dr < dmax(cut, lub_R)) {
                    Yc11 = 2.0 / 5.0; Yh11 = -1.0 / 10.0; h_c = dmax(dr, lub_R); beta = alpha * std::log(h_c / h);
                    Yc11 *= 8.0 * beta * rad * rad; Yh11 *= 8.0 * beta * rad * rad;
                } else Yc11 = Yh11 = 0.0;
                x12 /= r12; y12 /= r12; z12 /= r12;
                const double um_r = umx * x12 + umy * y12 + umz * z12;
                const double w1_r = w1x * x12 + w1y * y12 + w1z * z12;
                const double umxrx = umy * z12 - umz * y12, umxry = umz * x12 - umx * z12, umxrz = umx * y12 - umy * x12;
                const double w1xrx = w1y * z12 - w1z * y12, w1xry = w1z * x12 - w1x * z12, w1xrz = w1x * y12 - w1y * x12;
                double Fx = -Xa * x12 * um_r - Ya * (umx - um_r * x12);
                double Fy = -Xa * y12 * um_r - Ya * (umy - um_r * y12);
                double Fz = -Xa * z12 * um_r - Ya * (umz - um_r * z12);
                Fx += -Yb1 * w1xrx; Fy += -Yb1 * w1xry; Fz += -Yb1 * w1xrz;
                double T1x = Yb1 * umxrx, T1y = Yb1 * umxry, T1z = Yb1 * umxrz;
                T1x += -Yc11 * (w1x - w1_r * x12); T1y += -Yc11 * (w1y - w1_r * y12); T1z += -Yc11 * (w1z - w1_r * z12);
                P.FLub += V3(Fx, Fy, Fz); P.MLub += V3(T1x, T1y, T1z); W.FLub -= V3(Fx, Fy, Fz);
                double S1[6];  // DEM.cpp:2623-2634 (the wall's own bracketing of the Xg1 term)
                S1[0] = Xg1 * (um_r * x12 * x12) + Yg1 * (x12 * umx + umx * x12);
                S1[1] = Xg1 * (um_r * y12 * y12) + Yg1 * (y12 * umy + umy * y12);
                S1[2] = Xg1 * (um_r * z12 * z12) + Yg1 * (z12 * umz + umz * z12);
                S1[3] = Xg1 * (um_r * y12 * z12) + Yg1 * (y12 * umz + umy * z12);
                S1[4] = Xg1 * (um_r * z12 * x12) + Yg1 * (z12 * umx + umz * x12);
                S1[5] = Xg1 * (um_r * x12 * y12) + Yg1 * (x12 * umy + umx * y12);
                S1[0] += -Yh11 * (x12 * w1xrx + w1xrx * x12);
                S1[1] += -Yh11 * (y12 * w1xry + w1xry * y12);
                S1[2] += -Yh11 * (z12 * w1xrz + w1xrz * z12);
                S1[3] += -Yh11 * (y12 * w1xrz + w1xry * z12);
                S1[4] += -Yh11 * (z12 * w1xrx + w1xrz * x12);
                S1[5] += -Yh11 * (x12 * w1xry + w1xrx * y12);
                P.SLub.addSym(S1);
            }
    }
    // tangential spring-dashpot with Coulomb cap, shared by the wall and fictional-lid contacts
    // (DEM.cpp:1962-2063, 2175-2254); xi is the per-particle wall history xiw.
    void wallLikeContact(Particle& P, double x1w, double y1w, double z1w, V3& Fout, V3& Mout, V3* felOut = nullptr,
                         V3* fvisOut = nullptr) {
        double felx = 0, fely = 0, felz = 0, fvisx = 0, fvisy = 0, fvisz = 0;  // DEM.cpp:2001-2079
        const double radw = P.r;
        const double r1w = std::sqrt(x1w * x1w + y1w * y1w + z1w * z1w);
        const double urxw = P.x1.x, uryw = P.x1.y, urzw = P.x1.z;
        const double delw = radw - r1w;
        double fnx = 0, fny = 0, fnz = 0, ftx = 0, fty = 0, ftz = 0, n1wx = 0, n1wy = 0, n1wz = 0;
        if (delw > 0.0) {
            n1wx = x1w / r1w; n1wy = y1w / r1w; n1wz = z1w / r1w;
            const double gn1w = urxw * n1wx + uryw * n1wy + urzw * n1wz;
            const double muw = prm.frictionCoefPart, k = prm.linearStiff;
            fnx = k * delw * n1wx - dampCoeff * gn1w * n1wx;
            fny = k * delw * n1wy - dampCoeff * gn1w * n1wy;
            fnz = k * delw * n1wz - dampCoeff * gn1w * n1wz;
            felx = k * delw * n1wx; fely = k * delw * n1wy; felz = k * delw * n1wz;
            fvisx = -dampCoeff * gn1w * n1wx; fvisy = -dampCoeff * gn1w * n1wy; fvisz = -dampCoeff * gn1w * n1wz;
            if (muw != 0.0) {
                const double o1wx = P.r * P.w1.x, o1wy = P.r * P.w1.y, o1wz = P.r * P.w1.z;
                const double v1wx = urxw + n1wy * o1wz - n1wz * o1wy - gn1w * n1wx;
                const double v1wy = uryw + n1wz * o1wx - n1wx * o1wz - gn1w * n1wy;
                const double v1wz = urzw + n1wx * o1wy - n1wy * o1wx - gn1w * n1wz;
                const double xi1w = P.xiw.x * n1wx + P.xiw.y * n1wy + P.xiw.z * n1wz;
                P.xiw.x -= xi1w * n1wx; P.xiw.y -= xi1w * n1wy; P.xiw.z -= xi1w * n1wz;
                const double ktw = 0.2 * k;
                const double dct = prm.kinVisc == 0.0 ? dampCoeff : 0.0;
                ftx = -ktw * P.xiw.x - dct * v1wx;
                fty = -ktw * P.xiw.y - dct * v1wy;
                ftz = -ktw * P.xiw.z - dct * v1wz;
                const double fnnw = std::sqrt(fnx * fnx + fny * fny + fnz * fnz);
                const double fttw = std::sqrt(ftx * ftx + fty * fty + ftz * ftz);
                const double friw = muw * fnnw;
                const double t1wx = ftx / fttw, t1wy = fty / fttw, t1wz = ftz / fttw;
                P.xiw.x += v1wx * deltat; P.xiw.y += v1wy * deltat; P.xiw.z += v1wz * deltat;
                if (friw < fttw) {
                    ftx = friw * t1wx; fty = friw * t1wy; ftz = friw * t1wz;
                    P.xiw.x = -(1 / ktw) * (ftx + dct * v1wx);
                    P.xiw.y = -(1 / ktw) * (fty + dct * v1wy);
                    P.xiw.z = -(1 / ktw) * (ftz + dct * v1wz);
                }
                felx += ftx; fely += fty; felz += ftz;
            }
        } else {
            P.xiw = V3();
        }
        if (felOut) *felOut = V3(felx, fely, felz);
        if (fvisOut) *fvisOut = V3(fvisx, fvisy, fvisz);
        const double dfx = fnx + ftx, dfy = fny + fty, dfz = fnz + ftz;
        const double dtx = fty * n1wz - ftz * n1wy, dty = ftz * n1wx - ftx * n1wz, dtz = ftx * n1wy - fty * n1wx;
        Fout = V3(dfx, dfy, dfz);
        Mout = V3(P.r * dtx, P.r * dty, P.r * dtz);
    }
    void wallContacts() {  // DEM.cpp:1946-2153
        for (auto& P : particles)
            for (auto& W : walls) {
                const double x1w = (P.x0.x - W.p.x) * W.n.x * W.n.x;
                const double y1w = (P.x0.y - W.p.y) * W.n.y * W.n.y;
                const double z1w = (P.x0.z - W.p.z) * W.n.z * W.n.z;
                V3 Fc, Mc;
                V3 fel, fvis;
                wallLikeContact(P, x1w, y1w, z1w, Fc, Mc, &fel, &fvis);
                P.FContact += Fc; W.FContact -= Fc; P.MContact += Mc;
                P.FElas = P.FElas + fel;  // DEM.cpp:2103-2138
                P.SElas.addArm(x1w, y1w, z1w, fel.x, fel.y, fel.z);
                P.SVisc.addArm(x1w, y1w, z1w, fvis.x, fvis.y, fvis.z);
                P.SContact.addArm(x1w, y1w, z1w, Fc.x, Fc.y, Fc.z);  // DEM.cpp:2110-2148
                contactStress.addArm(x1w, y1w, z1w, Fc.x, Fc.y, Fc.z);
            }
    }
    void fictionalWallContacts() {  // DEM.cpp:2157-2296
        for (auto& P : particles) {
            V3 Fc, Mc;
            wallLikeContact(P, 0.0, P.x0.y - (demSize[1] - 0.3), 0.0, Fc, Mc);
            P.FContact += Fc; P.MContact += Mc;
        }
    }
    void objectParticleContacts() {  // DEM.cpp:912-1138
        for (auto& P : particles)
            for (auto& O : objects) {
                const double radw = 0.5 * (P.r + O.r);
                const double x1w = P.x0.x - O.x0.x, y1w = P.x0.y - O.x0.y, z1w = P.x0.z - O.x0.z;
                const double r1w = std::sqrt(x1w * x1w + y1w * y1w + z1w * z1w);
                const double n1wx = x1w / r1w, n1wy = y1w / r1w, n1wz = z1w / r1w;
                const double urxw = P.x1.x - O.x1.x, uryw = P.x1.y - O.x1.y, urzw = P.x1.z - O.x1.z;
                const double delw = (2 * radw) - r1w;
                double fnx = 0, fny = 0, fnz = 0, ftx = 0, fty = 0, ftz = 0;
                double felx = 0, fely = 0, felz = 0, fvisx = 0, fvisy = 0, fvisz = 0;  // DEM.cpp:970-1056
                if (delw > 0.0) {
                    const double gn1w = urxw * n1wx + uryw * n1wy + urzw * n1wz;
                    const double muw = prm.frictionCoefPart, k = prm.linearStiff;
                    fnx = k * delw * n1wx - dampCoeff * gn1w * n1wx;
                    fny = k * delw * n1wy - dampCoeff * gn1w * n1wy;
                    fnz = k * delw * n1wz - dampCoeff * gn1w * n1wz;
                    felx = k * delw * n1wx; fely = k * delw * n1wy; felz = k * delw * n1wz;
                    fvisx = -dampCoeff * gn1w * n1wx; fvisy = -dampCoeff * gn1w * n1wy; fvisz = -dampCoeff * gn1w * n1wz;
                    if (muw != 0.0) {
                        const double o1wx = P.r * P.w1.x + O.r * O.w1.x;
                        const double o1wy = P.r * P.w1.y + O.r * O.w1.y;
                        const double o1wz = P.r * P.w1.z + O.r * O.w1.z;
                        const double v1wx = urxw + n1wy * o1wz - n1wz * o1wy - gn1w * n1wx;
                        const double v1wy = uryw + n1wz * o1wx - n1wx * o1wz - gn1w * n1wy;
                        const double v1wz = urzw + n1wx * o1wy - n1wy * o1wx - gn1w * n1wz;
                        const double xi1w = P.xio.x * n1wx + P.xio.y * n1wy + P.xio.z * n1wz;
                        P.xio.x -= xi1w * n1wx; P.xio.y -= xi1w * n1wy; P.xio.z -= xi1w * n1wz;
                        const double ktw = 0.2 * k;
                        const double dct = 0.0;  // dampCoeffTanw, zero-initialised (see header)
                        ftx = -ktw * P.xio.x; fty = -ktw * P.xio.y; ftz = -ktw * P.xio.z;
                        const double fnnw = std::sqrt(fnx * fnx + fny * fny + fnz * fnz);
                        const double fttw = std::sqrt(ftx * ftx + fty * fty + ftz * ftz);
                        const double friw = muw * fnnw;
                        const double t1wx = ftx / fttw, t1wy = fty / fttw, t1wz = ftz / fttw;
                        if (friw < fttw) {
                            ftx = friw * t1wx; fty = friw * t1wy; ftz = friw * t1wz;
                            // the capped split as written (DEM.cpp:1000-1007): the z term of fel
                            // goes to y, and all three terms of fvis go to x
                            const double fne = std::sqrt(felx * felx + fely * fely + felz * felz);
                            const double fnv = std::sqrt(fvisx * fvisx + fvisy * fvisy + fvisz * fvisz);
                            felx += muw * fne * t1wx; fely += muw * fne * t1wy; fely += muw * fne * t1wz;
                            fvisx += muw * fnv * t1wx; fvisx += muw * fnv * t1wy; fvisx += muw * fnv * t1wz;
                            P.xio.x = -(1 / ktw) * (ftx + dct * v1wx);
                            P.xio.y = -(1 / ktw) * (fty + dct * v1wy);
                            P.xio.z = -(1 / ktw) * (ftz + dct * v1wz);
                        } else {
                            felx += -ktw * P.xio.x; fely += -ktw * P.xio.y; felz += -ktw * P.xio.z;
                            fvisx += -dct * v1wx; fvisy += -dct * v1wy; fvisz += -dct * v1wz;
                            P.xio.x += v1wx * deltat; P.xio.y += v1wy * deltat; P.xio.z += v1wz * deltat;
                        }
                    }
                } else {
                    P.xio = V3();
                }
                const double dfx = fnx + ftx, dfy = fny + fty, dfz = fnz + ftz;
                const double dtx = fty * n1wz - ftz * n1wy, dty = ftz * n1wx - ftx * n1wz, dtz = ftx * n1wy - fty * n1wx;
                P.FContact += V3(dfx, dfy, dfz);
                O.FContact -= V3(dfx, dfy, dfz);
                P.MContact += V3(P.r * dtx, P.r * dty, P.r * dtz);
                O.MContact += V3(O.r * dtx, O.r * dty, O.r * dtz);
                P.SContact.addArm(x1w, y1w, z1w, dfx, dfy, dfz);  // DEM.cpp:1097-1135
                P.FElas = P.FElas + V3(felx, fely, felz);
                P.SElas.addArm(x1w, y1w, z1w, felx, fely, felz);
                P.SVisc.addArm(x1w, y1w, z1w, fvisx, fvisy, fvisz);
                contactStress.addArm(x1w, y1w, z1w, dfx, dfy, dfz);
            }
    }
    void objectParticleLubrication() {  // DEM.cpp:1140-1340
        for (auto& P : particles)
            for (auto& O : objects) {
                bool act;
                LubOut o = pairLub(P.x0, O.x0, P.r, O.r, P.x1, O.x1, P.w1, O.w1, O.r, P.r, act);
                if (!act) continue;
                P.FLub += o.F; O.FLub -= o.F; P.MLub += o.T1; O.MLub += o.T2;
                P.SLub.addSym(o.S1);
            }
    }
    void particleContacts() {  // DEM.cpp:1487-1791
        const double k = prm.linearStiff, mu = prm.frictionCoefPart;
        for (int n1 = 0; n1 < int(particles.size()); n1++) {
            Particle& A = particles[n1];
            for (int nn = 1; nn <= A.list[0]; nn++) {
                const int n2 = A.list[nn];
                if (n2 >= n1) continue;
                Particle& B = particles[n2];
                const double rad = 0.5 * (A.r + B.r);
                const double x12 = A.x0.x - B.x0.x, y12 = A.x0.y - B.x0.y, z12 = A.x0.z - B.x0.z;
                const double r12 = std::sqrt(x12 * x12 + y12 * y12 + z12 * z12);
                const double urx = A.x1.x - B.x1.x, ury = A.x1.y - B.x1.y, urz = A.x1.z - B.x1.z;
                const double n12x = x12 / r12, n12y = y12 / r12, n12z = z12 / r12;
                const double gn12 = urx * n12x + ury * n12y + urz * n12z;
                const double del = 2 * rad - r12;
                double fnx = 0, fny = 0, fnz = 0, ftx = 0, fty = 0, ftz = 0;
                double felx = 0, fely = 0, felz = 0, fvisx = 0, fvisy = 0, fvisz = 0;  // DEM.cpp:1555-1679
                if (del > 0.0) {
                    fnx = k * del * n12x - dampCoeff * gn12 * n12x;
                    fny = k * del * n12y - dampCoeff * gn12 * n12y;
                    fnz = k * del * n12z - dampCoeff * gn12 * n12z;
                    felx = k * del * n12x; fely = k * del * n12y; felz = k * del * n12z;
                    fvisx = -dampCoeff * gn12 * n12x; fvisy = -dampCoeff * gn12 * n12y; fvisz = -dampCoeff * gn12 * n12z;
                    if (mu != 0.0) {
                        const double o12x = A.r * A.w1.x + B.r * B.w1.x;
                        const double o12y = A.r * A.w1.y + B.r * B.w1.y;
                        const double o12z = A.r * A.w1.z + B.r * B.w1.z;
                        const double v12x = urx + n12y * o12z - n12z * o12y - gn12 * n12x;
                        const double v12y = ury + n12z * o12x - n12x * o12z - gn12 * n12y;
                        const double v12z = urz + n12x * o12y - n12y * o12x - gn12 * n12z;
                        const double xi12 = A.xix[nn] * n12x + A.xiy[nn] * n12y + A.xiz[nn] * n12z;
                        A.xix[nn] -= xi12 * n12x; A.xiy[nn] -= xi12 * n12y; A.xiz[nn] -= xi12 * n12z;
                        const double kt = 0.2 * k;
                        const double dct = prm.kinVisc == 0.0 ? dampCoeff : 0.0;
                        ftx = -kt * A.xix[nn] - dct * v12x;
                        fty = -kt * A.xiy[nn] - dct * v12y;
                        ftz = -kt * A.xiz[nn] - dct * v12z;
                        const double fnn = std::sqrt(fnx * fnx + fny * fny + fnz * fnz);
                        const double ftt = std::sqrt(ftx * ftx + fty * fty + ftz * ftz);
                        const double fri = mu * fnn;
                        const double t12x = ftx / ftt, t12y = fty / ftt, t12z = ftz / ftt;
                        A.xix[nn] += v12x * deltat; A.xiy[nn] += v12y * deltat; A.xiz[nn] += v12z * deltat;
                        if (fri < ftt) {
                            ftx = fri * t12x; fty = fri * t12y; ftz = fri * t12z;
                            A.xix[nn] = -(1 / kt) * (ftx + dct * v12x);
                            A.xiy[nn] = -(1 / kt) * (fty + dct * v12y);
                            A.xiz[nn] = -(1 / kt) * (ftz + dct * v12z);
                        }
                        felx += ftx; fely += fty; felz += ftz;
                    }
                    // a contact counts where the box-wrapped centres are within 2 maxPartRadius
                    // (DEM.cpp:1647-1670)
                    const double x1 = dbox(A.x0.x, demSize[0]), y1 = dbox(A.x0.y, demSize[1]), z1 = dbox(A.x0.z, demSize[2]);
                    const double x2 = dbox(B.x0.x, demSize[0]), y2 = dbox(B.x0.y, demSize[1]), z2 = dbox(B.x0.z, demSize[2]);
                    const double zij = z1 - z2, xij = x1 - x2, yij = y1 - y2;
                    const double rij = std::sqrt(zij * zij + yij * yij + xij * xij);
                    if (rij <= 2.0 * (maxPartRadius)) { A.coord += 1.0; B.coord += 1.0; }
                } else {
                    A.xix[nn] = A.xiy[nn] = A.xiz[nn] = 0.0;
                }
                const double dfx = fnx + ftx, dfy = fny + fty, dfz = fnz + ftz;
                const double dtx = fty * n12z - ftz * n12y, dty = ftz * n12x - ftx * n12z, dtz = ftx * n12y - fty * n12x;
                A.FContact += V3(dfx, dfy, dfz);
                B.FContact -= V3(dfx, dfy, dfz);
                A.MContact += V3(A.r * dtx, A.r * dty, A.r * dtz);
                B.MContact += V3(B.r * dtx, B.r * dty, B.r * dtz);
                A.SContact.addArm(x12, y12, z12, dfx, dfy, dfz);  // DEM.cpp:1715-1783
                B.SContact.addArm(x12, y12, z12, dfx, dfy, dfz);
                A.FElas = A.FElas + V3(felx, fely, felz);
                B.FElas = B.FElas - V3(felx, fely, felz);
                A.SElas.addArm(x12, y12, z12, felx, fely, felz); A.SVisc.addArm(x12, y12, z12, fvisx, fvisy, fvisz);
                B.SElas.addArm(x12, y12, z12, felx, fely, felz); B.SVisc.addArm(x12, y12, z12, fvisx, fvisy, fvisz);
                contactStress.addArm(x12, y12, z12, dfx, dfy, dfz);
            }
        }
    }
    void particleElectro() {  // DEM.cpp:1404-1484
        const double fer = prm.electroForce, lamda = prm.debyeLength;
        for (int n1 = 0; n1 < int(particles.size()); n1++) {
            Particle& A = particles[n1];
            for (int nn = 1; nn <= A.list[0]; nn++) {
                const int n2 = A.list[nn];
                if (n2 >= n1) continue;
                Particle& B = particles[n2];
                const double rad = 0.5 * (A.r + B.r);
                const double x12 = A.x0.x - B.x0.x, y12 = A.x0.y - B.x0.y, z12 = A.x0.z - B.x0.z;
                const double r12 = std::sqrt(x12 * x12 + y12 * y12 + z12 * z12);
                const double n12x = x12 / r12, n12y = y12 / r12, n12z = z12 / r12;
                const double del = 2 * rad - r12;
                double frx, fry, frz;
                if (del <= 0.0) {
                    frx = fer * std::exp(del / lamda) * n12x;
                    fry = fer * std::exp(del / lamda) * n12y;
                    frz = fer * std::exp(del / lamda) * n12z;
                } else {
                    frx = fer * n12x; fry = fer * n12y; frz = fer * n12z;
                }
                A.FRep += V3(frx, fry, frz);
                B.FRep -= V3(frx, fry, frz);
                A.SRep.addArm(x12, y12, z12, frx, fry, frz);  // DEM.cpp:1451-1479
                B.SRep.addArm(x12, y12, z12, frx, fry, frz);
                electroStress.addArm(x12, y12, z12, frx, fry, frz);
            }
        }
    }
    void evaluateForces() {  // DEM.cpp:655-796
        totalStress.reset(); contactStress.reset(); lubStress.reset(); electroStress.reset(); hydroStress.reset();
        for (auto& P : particles) {
            P.FTotal.reset(); P.MTotal.reset(); P.FContact.reset(); P.MContact.reset();
            P.FLub.reset(); P.MLub.reset(); P.FRep.reset();
            P.SLub.reset(); P.SContact.reset(); P.SRep.reset(); P.STotal.reset();
            P.FElas.reset(); P.SElas.reset(); P.SVisc.reset(); P.SVisc_h.reset(); P.coord = 0.0; P.lub_coord = 0.0;
        }
        for (auto& O : objects) {
            O.FTotal.reset(); O.FContact.reset(); O.FLub.reset(); O.FSpring.reset();
            O.MTotal.reset(); O.MContact.reset(); O.MLub.reset();
        }
        for (auto& W : walls) { W.FTotal.reset(); W.FContact.reset(); W.FLub.reset(); W.FRep.reset(); }
        for (auto& B : bodies) { B.FTotal.reset(); B.FSpring.reset(); }
        if (int(demTimeStep) >= prm.objWaitTime) {
            if (!bodies.empty()) bodyObjectInteraction();
            particleLubrication();
            wallLubrication();
            objectParticleContacts();
            objectParticleLubrication();
        } else {
            fictionalWallContacts();
        }
        particleContacts();
        particleElectro();
        wallContacts();
        for (auto& P : particles) {  // DEM.cpp:745-764
            for (int e = 0; e < 9; ++e) {
                lubStress.m[e] += P.SLub.m[e];
                hydroStress.m[e] += P.SHydro.m[e];
                P.STotal.m[e] += P.SHydro.m[e] + P.SLub.m[e] + P.SContact.m[e] + P.SRep.m[e];
                P.SVisc_h.m[e] += P.SHydro.m[e] + P.SLub.m[e];
                P.SVisc.m[e] += P.SVisc_h.m[e];
            }
            P.overlap = P.FElas / prm.linearStiff;  // DEM.cpp:761
            P.FTotal += P.FHydro + P.FContact + P.FLub + P.FGrav + P.FRep;
            P.MTotal += P.MHydro + P.MContact + P.MLub;
        }
        for (int e = 0; e < 9; ++e) {  // DEM.cpp:766-771
            lubStress.m[e] /= vol; contactStress.m[e] /= vol; electroStress.m[e] /= vol; hydroStress.m[e] /= vol;
            totalStress.m[e] += contactStress.m[e] + lubStress.m[e] + electroStress.m[e] + hydroStress.m[e];
        }
        for (auto& W : walls) W.FTotal += W.FHydro + W.FContact + W.FLub + W.FRep;
        if (int(demTimeStep) <= prm.objWaitTime)
            for (auto& P : particles) P.FGrav = V3();
        for (auto& O : objects) {
            O.FTotal += O.FHydro + O.FContact + O.FLub + O.FGrav + O.FSpring;
            O.MTotal += O.MHydro + O.MContact + O.MLub;
        }
        for (auto& B : bodies) B.FTotal += B.FGrav + B.FSpring;
    }
    // DEM.cpp:798-869: every body pulls every impactor along y with a spring of rest length L0
    // (L = the y distance), with dashpots and cubic / Lennard-Jones / WCA terms by nonLinType
    void bodyObjectInteraction() {
        const double springDamp = prm.springDamp, k13 = prm.k13, epsilon = prm.epsilon;
        for (auto& B : bodies)
            for (auto& O : objects) {
                double fs = 0.0, fdo = 0.0, fdb = 0.0;
                B.L = B.x0.y - O.x0.y;
                const V3 r = B.x0 - O.x0;
                const V3 n = V3(r.x / r.norm(), r.y / r.norm(), r.z / r.norm());
                const double del = B.L0 - B.L;
                const double bodyDamp = 2.0 * std::sqrt(B.m * B.k) * springDamp;
                const double objDamp = 2.0 * std::sqrt(O.m * B.k) * springDamp;
                const double sigma = B.r;
                switch (prm.nonLinType) {
                    case 0: fs = B.k * del * n.y; break;
                    case 1: fs = (B.k * del) * n.y; fdo = -O.x1.y * objDamp; fdb = -B.x1.y * bodyDamp; break;
                    case 2:
                        fs = (B.k * del + k13 * B.k * (del * del * del)) * n.y;
                        fdo = -O.x1.y * objDamp; fdb = -B.x1.y * bodyDamp;
                        break;
                    case 3:
                        fs = (B.k * del + (24.0 * epsilon / r.norm()) * 2.0 * (std::pow((sigma / r.norm()), 12.0))) * n.y;
                        fdo = -O.x1.y * objDamp; fdb = -B.x1.y * bodyDamp;
                        break;
                    default: {
                        const double wcacut = std::pow(2.0, 1.0 / 6.0) * sigma;
                        if (r.norm() <= wcacut)
                            fs = (B.k * del + (24.0 * epsilon / r.norm()) *
                                                  ((2.0 * std::pow(sigma / r.norm(), 12.0)) - (std::pow(sigma / r.norm(), 6.0)))) * n.y;
                        else
                            fs = (B.k * del) * n.y;
                        fdo = -O.x1.y * objDamp; fdb = -B.x1.y * bodyDamp;
                        break;
                    }
                }
                O.FSpring.y -= fs;
                B.FSpring.y += fs;
                O.FSpring.y += fdo;
                B.FSpring.y += fdb;
            }
    }
    void integration() {  // DEM.cpp:2676-2765
        for (auto& P : particles) {
            P.x0.x += deltat * P.x1.x + deltat * deltat * P.FTotal.x * P._m;
            P.x0.y += deltat * P.x1.y + deltat * deltat * P.FTotal.y * P._m;
            P.x0.z += deltat * P.x1.z + deltat * deltat * P.FTotal.z * P._m;
            P.x1.x += deltat * P.FTotal.x * P._m;
            P.x1.y += deltat * P.FTotal.y * P._m;
            P.x1.z += deltat * P.FTotal.z * P._m;
            P.w1.x += deltat * P.MTotal.x * P._I;
            P.w1.y += deltat * P.MTotal.y * P._I;
            P.w1.z += deltat * P.MTotal.z * P._I;
            P.disp = P.x_i - P.x0;  // DEM.cpp:2705
        }
        if (int(demTimeStep) == prm.objWaitTime)  // DEM.cpp:2721-2725
            for (auto& P : particles) P.x_i = P.x0;
        if (int(demTimeStep) >= prm.objWaitTime)
            for (auto& B : bodies) {  // DEM.cpp:2730-2740
                B.x0.x += deltat * B.x1.x + deltat * deltat * B.FTotal.x * B._m;
                B.x0.y += deltat * B.x1.y + deltat * deltat * B.FTotal.y * B._m;
                B.x0.z += deltat * B.x1.z + deltat * deltat * B.FTotal.z * B._m;
                B.x1.x += deltat * B.FTotal.x * B._m;
                B.x1.y += deltat * B.FTotal.y * B._m;
                B.x1.z += deltat * B.FTotal.z * B._m;
            }
        if (int(demTimeStep) >= prm.objWaitTime)
            for (auto& O : objects) {
                O.x0.x += deltat * O.x1.x + deltat * deltat * O.FTotal.x * O._m;
                O.x0.y += deltat * O.x1.y + deltat * deltat * O.FTotal.y * O._m;
                O.x0.z += deltat * O.x1.z + deltat * deltat * O.FTotal.z * O._m;
                O.x1.x += deltat * O.FTotal.x * O._m;
                O.x1.y += deltat * O.FTotal.y * O._m;
                O.x1.z += deltat * O.FTotal.z * O._m;
                O.w1.x += deltat * O.MTotal.x * O._I;
                O.w1.y += deltat * O.MTotal.y * O._I;
                O.w1.z += deltat * O.MTotal.z * O._I;
            }
    }
    void pbcs() {  // DEM.cpp:385-398, 2769-2836 (loop index pinned to start at 0)
        double* sz = demSize;
        auto wrap = [&](int a) {
            for (auto& P : particles) {
                double& c = a == 0 ? P.x0.x : (a == 1 ? P.x0.y : P.x0.z);
                if (c > sz[a]) c -= sz[a];
                if (c < 0.0) c += sz[a];
            }
        };
        if (boundary[0] == 4 && boundary[1] == 4) wrap(0);
        if (boundary[2] == 4 && boundary[3] == 4) wrap(1);
        if (boundary[4] == 4 && boundary[5] == 4) wrap(2);
    }
    int demStep() {  // DEM.cpp:359-401
        demTimeStep++;
        demTime += deltat;
        double maxVel = 0.0;  // evalMaxDisp (DEM.cpp:460-470)
        for (auto& P : particles) if (P.x1.norm2() > maxVel) maxVel = P.x1.norm2();
        maxVel = std::sqrt(maxVel);
        maxDisp += maxVel * deltat;
        int rc = 0;
        if (maxDisp > 0.05) { maxDisp = 0.0; rc = neighborList(); newNeighborList = true; }
        evaluateForces();
        integration();
        pbcs();
        return rc;
    }
    // goCycle (main.cpp:38-67) without IO::outputStep
    int cycleOnce() {
        realTime += unit.Time;
        ++currentTimeStep;
        ++lbTime;
        int rc = 0;
        if (prm.demSolve) { rc = demStep(); if (rc) return rc; }
        if (demTime >= demInitialRepeat) {
            if (prm.demSolve) couplingStep();
            rc = lbStep(); if (rc) return rc;
            if (freeSurface) rc = fsStep();
        } else if (demTime <= demInitialRepeat) {
            if (prm.demSolve) couplingStep();
        }
        return rc;
    }
};

extern "C" {

int lbdem_create(const lbdem_params* params, lbdem_ctx** out) {
    if (!params || !out) return LBDEM_E_ARG;
    lbdem_ctx* c = new lbdem_ctx();
    c->prm = *params;
    int rc = c->getParams();
    if (rc) { *out = c; return rc; }
    *out = c;
    return 0;
}
int lbdem_destroy(lbdem_ctx* ctx) { delete ctx; return 0; }
const char* lbdem_last_error(const lbdem_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lbdem_set_particles(lbdem_ctx* c, int n, const int* size, const double* r, const double* x0) {
    if (!c || n < 0 || (n > 0 && (!r || !x0))) return LBDEM_E_ARG;
    if (n > 65535) return c->fail(LBDEM_E_ARG, "more than 65535 particles (solidIndex is 16-bit, node.h:135)");
    c->particles.assign(n, Particle());
    for (int i = 0; i < n; ++i) {
        c->particles[i].size = size ? unsigned(size[i]) : 1u;
        c->particles[i].r = r[i];
        c->particles[i].x0 = V3(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]);
    }
    return 0;
}
int lbdem_set_objects(lbdem_ctx* c, int n, const double* r, const double* x0, const double* x1) {
    if (!c || n < 0 || (n > 0 && (!r || !x0 || !x1))) return LBDEM_E_ARG;
    c->objects.assign(n, Object());
    for (int i = 0; i < n; ++i) {
        c->objects[i].index = unsigned(i);
        c->objects[i].r = r[i];
        c->objects[i].x0 = V3(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]);
        c->objects[i].x1 = V3(x1[3 * i], x1[3 * i + 1], x1[3 * i + 2]);
    }
    return 0;
}
int lbdem_set_bodies(lbdem_ctx* c, int n, const double* r, const double* x0, const double* x1, const double* k,
                     const double* L0) {
    if (!c || n < 0 || n > LBDEM_MAX_BODIES || (n > 0 && (!r || !x0 || !x1 || !k || !L0))) return LBDEM_E_ARG;
    if (n > 0 && (c->prm.nonLinType < 0 || c->prm.nonLinType > 4))
        return c->fail(LBDEM_E_ARG, "nonLinType must be 0-4 with bodies");
    c->bodies.assign(n, Body());
    for (int i = 0; i < n; ++i) {
        Body& B = c->bodies[i];
        B.r = r[i];
        B.x0 = V3(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]);
        B.x1 = V3(x1[3 * i], x1[3 * i + 1], x1[3 * i + 2]);
        B.k = k[i];
        B.L0 = L0[i];
    }
    return 0;
}
// the fluidRecycle file's contents (LB.cpp:352-367): kept for latticeBolzmannInit
int lbdem_set_restart_fluid(lbdem_ctx* c, const int size[3], const uint8_t* types, int nRecords, const double* records) {
    if (!c || !size || !types || nRecords < 0 || (nRecords > 0 && !records)) return LBDEM_E_ARG;
    if (c->initialised) return c->fail(LBDEM_E_STATE, "the restart must be set before lbdem_init");
    for (int a = 0; a < 3; ++a)
        if (unsigned(size[a]) != c->lbSize[a]) return c->fail(LBDEM_E_ARG, "the restart file's lattice size differs");
    const size_t n = size_t(c->lbSize[0]) * c->lbSize[1] * c->lbSize[2];
    c->restartTypes.assign(types, types + n);
    c->restartRecords.assign(records, records + size_t(nRecords) * LBDEM_RESTART_REC);
    c->lbRestart = true;
    return 0;
}
int lbdem_init(lbdem_ctx* c) {
    if (!c) return LBDEM_E_ARG;
    c->latticeDefinition();
    c->discreteElementInit();
    c->latticeBolzmannInit();
    if (c->restartShort) return c->fail(LBDEM_E_ARG, "the restart file holds fewer records than active nodes");
    c->lbTime = 0; c->demTime = 0.0; c->realTime = 0.0; c->demTimeStep = 0;
    c->initialised = true;
    return c->err.empty() ? 0 : LBDEM_E_CAPACITY;
}
#define NEED_INIT if (!c || !c->initialised) return LBDEM_E_STATE
int lbdem_dem_step(lbdem_ctx* c) { NEED_INIT; return c->demStep(); }
int lbdem_coupling_step(lbdem_ctx* c) { NEED_INIT; c->couplingStep(); return 0; }
int lbdem_lb_step(lbdem_ctx* c) { NEED_INIT; return c->lbStep(); }
int lbdem_fs_step(lbdem_ctx* c) { NEED_INIT; return c->fsStep(); }
int lbdem_cycle(lbdem_ctx* c, int nsteps) {
    NEED_INIT;
    for (int s = 0; s < nsteps; ++s) { int rc = c->cycleOnce(); if (rc) return rc; }
    return 0;
}
int lbdem_synchronize(lbdem_ctx* c) { return c ? 0 : LBDEM_E_ARG; }

int lbdem_get_info(lbdem_ctx* c, lbdem_info* in) {
    NEED_INIT;
    std::memset(in, 0, sizeof(*in));
    for (int a = 0; a < 3; ++a) in->size[a] = int(c->lbSize[a]);
    in->totNodes = c->N;
    in->nParticles = int(c->particles.size()); in->nObjects = int(c->objects.size()); in->nWalls = int(c->walls.size());
    int na = 0, nf = 0, ni = 0, np = 0, no = 0;
    for (int64_t i = 0; i < c->N; ++i) {
        if (c->t[i] == 0) nf++;
        if (c->t[i] == 3) ni++;
        if (c->t[i] == 12) no++;
        if (c->p[i]) np++;
    }
    na = nf + ni;
    in->nActive = na; in->nFluid = nf; in->nInterface = ni; in->nParticleNodes = np; in->nObjectNodes = no;
    in->lbTime = c->lbTime; in->demTimeStep = c->demTimeStep; in->demTime = c->demTime; in->realTime = c->realTime;
    in->neighborListBuilds = c->nlBuilds; in->newNeighborList = c->newNeighborList ? 1 : 0;
    in->totalMassInit = c->totalMass; in->lastExtraMass = c->lastExtraMass; in->lastSurplus = c->lastSurplus;
    in->precision = LBDEM_PREC_FP64;
    in->nBodies = int(c->bodies.size());
    return 0;
}
int lbdem_get_types(lbdem_ctx* c, uint8_t* type, uint8_t* pflag, int32_t* solid) {
    NEED_INIT;
    for (int64_t i = 0; i < c->N; ++i) {
        if (type) type[i] = c->t[i];
        if (pflag) pflag[i] = c->p[i];
        if (solid) solid[i] = c->solid[i];
    }
    return 0;
}
int lbdem_get_fields(lbdem_ctx* c, double* f, double* mass, double* n) {
    NEED_INIT;
    const int64_t N = c->N;
    for (int64_t i = 0; i < N; ++i) {
        const bool e = c->exists[i];
        if (f) for (int q = 0; q < Q; ++q) f[q * N + i] = e ? c->f[size_t(i) * Q + q] : 0.0;
        if (mass) mass[i] = e ? c->mass[i] : 0.0;
        if (n) n[i] = e ? c->n[i] : 0.0;
    }
    return 0;
}
// node->u of every active node (IO.cpp:973-997 reads it); the oracle keeps it for every node
int lbdem_store_velocity(lbdem_ctx* c) { NEED_INIT; return 0; }
int lbdem_get_velocity(lbdem_ctx* c, double* uo, double* vo) {
    NEED_INIT;
    for (int64_t i = 0; i < c->N; ++i) {
        const bool a = c->exists[i] && c->isActive(i);
        if (uo) for (int k = 0; k < 3; ++k) uo[3 * i + k] = a ? c->u[3 * i + k] : 0.0;
        if (vo) vo[i] = a ? c->visc[i] : 0.0;
    }
    return 0;
}
// the CPU oracle allocates nothing per call either way: the same values as lbdem_get_fields
int lbdem_get_fields_noalloc(lbdem_ctx* c, double* f, double* mass, double* n) { return lbdem_get_fields(c, f, mass, n); }
static void put3(double* dst, const V3& v) { dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; }
int lbdem_get_particles(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    for (size_t i = 0; i < c->particles.size(); ++i) {
        const Particle& P = c->particles[i];
        double* r = rec + i * LBDEM_P_REC;
        put3(r + LBDEM_P_X0, P.x0); put3(r + LBDEM_P_X1, P.x1); put3(r + LBDEM_P_W1, P.w1);
        put3(r + LBDEM_P_FHYDRO, P.FHydro); put3(r + LBDEM_P_MHYDRO, P.MHydro);
        put3(r + LBDEM_P_FCONTACT, P.FContact); put3(r + LBDEM_P_MCONTACT, P.MContact);
        put3(r + LBDEM_P_FLUB, P.FLub); put3(r + LBDEM_P_MLUB, P.MLub); put3(r + LBDEM_P_FREP, P.FRep);
        put3(r + LBDEM_P_FTOTAL, P.FTotal); put3(r + LBDEM_P_MTOTAL, P.MTotal); put3(r + LBDEM_P_FGRAV, P.FGrav);
        r[LBDEM_P_FLUIDVOL] = P.fluidVolume;
        put3(r + LBDEM_P_XIW, P.xiw); put3(r + LBDEM_P_XIO, P.xio);
        r[LBDEM_P_R] = P.r; r[LBDEM_P_M] = P.m;
    }
    return 0;
}
int lbdem_get_objects(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    for (size_t i = 0; i < c->objects.size(); ++i) {
        const Object& O = c->objects[i];
        double* r = rec + i * LBDEM_O_REC;
        put3(r + LBDEM_O_X0, O.x0); put3(r + LBDEM_O_X1, O.x1); put3(r + LBDEM_O_W1, O.w1);
        put3(r + LBDEM_O_FHYDRO, O.FHydro); put3(r + LBDEM_O_MHYDRO, O.MHydro);
        put3(r + LBDEM_O_FCONTACT, O.FContact); put3(r + LBDEM_O_MCONTACT, O.MContact);
        put3(r + LBDEM_O_FLUB, O.FLub); put3(r + LBDEM_O_MLUB, O.MLub);
        put3(r + LBDEM_O_FTOTAL, O.FTotal); put3(r + LBDEM_O_MTOTAL, O.MTotal); put3(r + LBDEM_O_FGRAV, O.FGrav);
        r[LBDEM_O_R] = O.r; r[LBDEM_O_M] = O.m;
    }
    return 0;
}
int lbdem_get_bodies(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    for (size_t i = 0; i < c->bodies.size(); ++i) {
        const Body& B = c->bodies[i];
        double* r = rec + i * LBDEM_B_REC;
        put3(r + LBDEM_B_X0, B.x0); put3(r + LBDEM_B_X1, B.x1); put3(r + LBDEM_B_FSPRING, B.FSpring);
        put3(r + LBDEM_B_FTOTAL, B.FTotal); put3(r + LBDEM_B_FGRAV, B.FGrav);
        r[LBDEM_B_L] = B.L; r[LBDEM_B_L0] = B.L0; r[LBDEM_B_K] = B.k; r[LBDEM_B_R] = B.r; r[LBDEM_B_M] = B.m;
    }
    return 0;
}
int lbdem_set_stress(lbdem_ctx* c, int) {  // the oracle computes the stresses on every step
    if (!c) return LBDEM_E_ARG;
    return c->initialised ? c->fail(LBDEM_E_STATE, "lbdem_set_stress after lbdem_init") : 0;
}
int lbdem_get_stress(lbdem_ctx* c, double* global, double* stotal, double* disp) {
    NEED_INIT;
    if (global) {
        const T9* t[5] = {&c->totalStress, &c->hydroStress, &c->lubStress, &c->contactStress, &c->electroStress};
        for (int k = 0; k < 5; ++k) for (int e = 0; e < 9; ++e) global[9 * k + e] = t[k]->m[e];
    }
    for (size_t i = 0; i < c->particles.size(); ++i) {
        const Particle& P = c->particles[i];
        if (stotal) for (int e = 0; e < 9; ++e) stotal[9 * i + e] = P.STotal.m[e];
        if (disp) put3(disp + 3 * i, P.disp);
    }
    return 0;
}
int lbdem_get_particle_stress(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    for (size_t i = 0; i < c->particles.size(); ++i) {
        const Particle& P = c->particles[i];
        double* R = rec + i * LBDEM_S_REC;
        for (int e = 0; e < 9; ++e) {
            R[LBDEM_S_STOTAL + e] = P.STotal.m[e]; R[LBDEM_S_SELAS + e] = P.SElas.m[e];
            R[LBDEM_S_SVISC + e] = P.SVisc.m[e]; R[LBDEM_S_SVISC_H + e] = P.SVisc_h.m[e];
        }
        put3(R + LBDEM_S_OVERLAP, P.overlap);
        put3(R + LBDEM_S_DISP, P.disp);
        R[LBDEM_S_COORD] = P.coord; R[LBDEM_S_LUB_COORD] = P.lub_coord;
    }
    return 0;
}
int lbdem_get_walls(lbdem_ctx* c, double* rec) {
    NEED_INIT;
    for (size_t i = 0; i < c->walls.size(); ++i) {
        const Wall& W = c->walls[i];
        double* r = rec + i * LBDEM_W_REC;
        put3(r + LBDEM_W_FHYDRO, W.FHydro); put3(r + LBDEM_W_FCONTACT, W.FContact);
        put3(r + LBDEM_W_FLUB, W.FLub); put3(r + LBDEM_W_FTOTAL, W.FTotal);
    }
    return 0;
}
int lbdem_get_neighbors(lbdem_ctx* c, int cap, int* count, int* list) {
    NEED_INIT;
    for (size_t i = 0; i < c->particles.size(); ++i) {
        const Particle& P = c->particles[i];
        std::vector<int> v(P.list + 1, P.list + 1 + P.list[0]);
        std::sort(v.begin(), v.end());
        count[i] = int(v.size());
        for (int k = 0; k < int(v.size()) && k < cap; ++k) list[i * cap + k] = v[k];
    }
    return 0;
}
int lbdem_profile_enable(lbdem_ctx* c, int) { NEED_INIT; return 0; }
int lbdem_profile_read(lbdem_ctx* c, double* fused_ms, int* n) {
    NEED_INIT;
    if (fused_ms) *fused_ms = 0.0;
    if (n) *n = 0;
    return 0;
}
// device memory, graph replay and the run pool belong to the device library: the oracle holds
// no device memory, steps one by one and has no pool
int lbdem_estimate_bytes(const lbdem_params*, int, const double*, int, const double*, size_t* bytes) {
    if (bytes) *bytes = 0;
    return LBDEM_E_STATE;
}
int lbdem_device_bytes(lbdem_ctx* c, size_t* bytes) {
    if (!c || !bytes) return LBDEM_E_ARG;
    *bytes = 0;
    return 0;
}
int lbdem_use_graphs(lbdem_ctx* c, int) { return c ? 0 : LBDEM_E_ARG; }
int lbdem_set_tile(lbdem_ctx* c, int tileZ) {
    if (!c) return LBDEM_E_ARG;
    return (tileZ == 0 || tileZ == 1 || tileZ == 2 || tileZ == 4 || tileZ == 8) ? 0 : LBDEM_E_ARG;
}
int lbdem_get_tuning(lbdem_ctx* c, lbdem_tuning* t) {
    if (!c || !t) return LBDEM_E_ARG;
    std::memset(t, 0, sizeof(*t));  // no kernel tiles on the CPU
    return 0;
}
int lbdem_batch_create(lbdem_ctx* const*, int, lbdem_batch** out) {
    if (out) *out = nullptr;
    return LBDEM_E_STATE;
}
int lbdem_batch_cycle(lbdem_batch*, int) { return LBDEM_E_STATE; }
int lbdem_batch_synchronize(lbdem_batch*) { return LBDEM_E_STATE; }
const char* lbdem_batch_last_error(const lbdem_batch*) { return "the CPU oracle has no batches"; }
int lbdem_batch_destroy(lbdem_batch*) { return 0; }
int lbdem_batch_set_tile(lbdem_batch*, int) { return LBDEM_E_STATE; }
int lbdem_batch_get_tile(const lbdem_batch*, int*, int*) { return LBDEM_E_STATE; }
int lbdem_batch_profile_enable(lbdem_batch*, int) { return LBDEM_E_STATE; }
int lbdem_batch_profile_read(const lbdem_batch*, int, const char** kernel, double*, int*, int* count) {
    if (kernel) *kernel = nullptr;
    if (count) *count = 0;
    return LBDEM_E_STATE;
}
int lbdem_pool_create(size_t, int, int, int, lbdem_pool** out) {
    if (out) *out = nullptr;
    return LBDEM_E_STATE;
}
int lbdem_pool_submit(lbdem_pool*, const lbdem_run_desc*, int*) { return LBDEM_E_STATE; }
int lbdem_pool_poll(lbdem_pool*, int, lbdem_run_result*) { return LBDEM_E_STATE; }
int lbdem_pool_wait(lbdem_pool*, int, double, lbdem_run_result*) { return LBDEM_E_STATE; }
int lbdem_pool_drain(lbdem_pool*, double) { return LBDEM_E_STATE; }
int lbdem_pool_cancel(lbdem_pool*, int) { return LBDEM_E_STATE; }
int lbdem_pool_get_stats(lbdem_pool*, lbdem_pool_stats*) { return LBDEM_E_STATE; }
int lbdem_pool_set_batching(lbdem_pool*, int) { return LBDEM_E_STATE; }
int lbdem_pool_set_batch_tile(lbdem_pool*, int) { return LBDEM_E_STATE; }
double lbdem_pool_clock(const lbdem_pool*) { return 0.0; }
const char* lbdem_pool_last_error(const lbdem_pool*) { return "the CPU oracle has no run pool"; }
int lbdem_pool_destroy(lbdem_pool*) { return 0; }

int lbdem_total_mass(lbdem_ctx* c, double* m) {
    NEED_INIT;
    double s = 0.0;
    for (int64_t i = 0; i < c->N; ++i)
        if ((c->t[i] == 0 || c->t[i] == 3) && !c->p[i] && c->exists[i]) s += c->mass[i];
    *m = s;
    return 0;
}

}  // extern "C"
