// Host facade of the coupled LB / DEM step with the reference's class and method names
// (LB.h:166-172, DEM.h:152-154, IO.h:63-64, main.cpp:38-67, 128-185). The solver state lives
// behind the C ABI (include/lbdem.h) on the device; the public members that the reference's IO
// reads are host mirrors refreshed by `refresh()` (a download on the output interval only).
#pragma once

#include <cstdint>
#include <iosfwd>
#include <string>
#include <vector>

#include "abi.hpp"
#include "config.hpp"

namespace lbdem {

constexpr int Q_DIRECTIONS = 19;  // D3Q19 (lattice.h)

struct tVect {
    double x = 0.0, y = 0.0, z = 0.0;
    tVect() = default;
    tVect(double a, double b, double c) : x(a), y(b), z(c) {}
    tVect operator+(const tVect& o) const { return {x + o.x, y + o.y, z + o.z}; }
    tVect operator-(const tVect& o) const { return {x - o.x, y - o.y, z - o.z}; }
    tVect operator*(double s) const { return {x * s, y * s, z * s}; }
    tVect operator/(double s) const { return {x / s, y / s, z / s}; }
    double dot(const tVect& o) const { return x * o.x + y * o.y + z * o.z; }
    // "x y z " with the stream's current formatting (the reference's tVect::print)
    void print(std::ostream& os) const;
    // "x y z\n" fixed with 8 digits, left set on the stream (tinyVector::printFixedLine, myvector.cpp:25-27)
    void printFixedLine(std::ostream& os) const;
};

// tinyMatrix (myvector.h) with the reference's three printers (myvector.cpp:382-392)
struct tMat {
    double m00 = 0, m01 = 0, m02 = 0, m10 = 0, m11 = 0, m12 = 0, m20 = 0, m21 = 0, m22 = 0;
    void set(const double* v) {
        m00 = v[0]; m01 = v[1]; m02 = v[2]; m10 = v[3]; m11 = v[4]; m12 = v[5]; m20 = v[6]; m21 = v[7]; m22 = v[8];
    }
    void print(std::ostream& os) const;        // "xx yy zz xy yz zx " (symmetrised shear)
    void printNormal(std::ostream& os) const;  // "xx yy zz "
    void printShear(std::ostream& os) const;   // "xy yz zx "
};

// host mirrors of the element records (fields on the path; elmt.h:15-228, utils.h:19-101)
struct particle {
    unsigned index = 0;
    int size = 1;
    double r = 0.0, m = 0.0, fluidVolume = 0.0;
    tVect x0, x1, w1, FHydro, MHydro, FContact, MContact, FLub, MLub, FRep, FTotal, MTotal, FGrav, xiw, xio;
    tMat STotal, SElas, SVisc, SVisc_h;  // with the stress diagnostics on
    tVect disp, overlap;
    double coord = 0.0, lub_coord = 0.0, pressure = 0.0, disp_r = 0.0;
};
struct object {
    unsigned index = 0;
    double r = 0.0, m = 0.0;
    tVect x0, x1, w1, FHydro, MHydro, FContact, MContact, FLub, MLub, FRep, FTotal, MTotal, FGrav;
};
struct wall {
    unsigned index = 0;
    tVect FHydro, FContact, FLub, FTotal;
};
struct body {
    unsigned index = 0;
    double r = 0.0, m = 0.0, k = 0.0, L0 = 0.0, L = 0.0;
    tVect x0, x1, FSpring, FTotal, FGrav;
};
using particleList = std::vector<particle>;
using objectList = std::vector<object>;
using bodyList = std::vector<body>;
using wallList = std::vector<wall>;
using unsIntList = std::vector<unsigned>;
using doubleList = std::vector<double>;
using vecList = std::vector<tVect>;

// One solver instance: the loaded ABI, the parameters gathered by the Get methods, the inputs
// read from input/*.dat and the device context once created.
class Context {
public:
    explicit Context(const std::string& libraryPath, int precision = LBDEM_PREC_FP64, int device = 0);
    ~Context();
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    const Api& api() const { return api_; }
    lbdem_params params{};
    std::vector<int> psize;
    std::vector<double> pr, px0, orad, ox0, ox1;
    std::vector<double> brad, bx0, bx1, bk, bL0;  // input/body.dat (DEM.cpp:168-196)
    bool stress = false;  // compute the stress diagnostics (lbdem_set_stress before lbdem_init)

    lbdem_ctx* ctx() const { return ctx_; }
    void create();                    // lbdem_create + inputs (idempotent)
    void initialise();                // lbdem_init (idempotent)
    bool initialised() const { return initialised_; }
    // status check with the reference's behaviour: message and exit(1) (macros.h:10)
    void check(int rc, const char* what) const;

private:
    Api api_;
    lbdem_ctx* ctx_ = nullptr;
    bool initialised_ = false;
};

class LB {
public:
    explicit LB(Context& c) : c_(c) {}
    // LB.cpp:128-338 (the keys of the Newtonian path)
    void latticeBoltzmannGet(ConfigFile& configFile, CommandLine& commandLine);
    void latticeDefinition();                                                       // LB.cpp:85-126
    void LBShow() const;
    void latticeBolzmannInit(wallList& walls, particleList& particles, objectList& objects);  // LB.cpp:340-398
    void latticeBolzmannStep(particleList& particles, wallList& walls, objectList& objects);  // LB.cpp:400-418
    void latticeBoltzmannCouplingStep(bool& newNeighborList, particleList& particles, objectList& objects);  // LB.cpp:432-467
    void latticeBoltzmannFreeSurfaceStep();                                           // LB.cpp:422-430
    // total fluid mass over active nodes outside particles (LB.cpp:388-395)
    double currentMass() const;
    // the next LB step keeps every active node's velocity for the output (IO reads node->u)
    void storeVelocity();

    unsigned int time = 0;
    int freeSurface = 0, forceField = 0, lbmInitialRepeat = 0;
    int lbRestart = 0;              // LB.cpp:138-139
    std::string lbRestartFile;
    double fluidHeight = 0.0;   // physical, as configured
    double unitTime = 1.0, unitLength = 1.0, unitDensity = 1.0;
    // composite units (measureUnits::setComposite, node.cpp:512-525)
    double unitSpeed() const { return unitLength / unitTime; }
    double unitPressure() const { return unitDensity * unitLength * unitLength / unitTime / unitTime; }
    unsigned lbSize[3] = {0, 0, 0};
    double totalMass = 0.0;     // at initialisation (LB.cpp:388-395)
    Context& context() const { return c_; }

private:
    Context& c_;
};

class DEM {
public:
    explicit DEM(Context& c) : c_(c) {}
    // DEM.cpp:17-197: parameters, input/particles.dat and input/objects.dat (relative to the
    // working directory, as the reference reads them)
    void discreteElementGet(ConfigFile& configFile, CommandLine& commandLine);
    // DEM.cpp:199-357 (arguments kept for signature parity; the device derives them)
    void discreteElementInit(const unsIntList& boundary, const doubleList& externalBoundary,
                             const vecList& boundaryLocation, const double& externalTimeStep,
                             const tVect externalForce, const double& fluidHeight, bool& freeSurface);
    void discreteElementStep(const unsIntList& boundary);                            // DEM.cpp:359-401
    // downloads particle, object and wall records into the mirrors below
    void refresh();

    double demTime = 0.0, demInitialRepeat = 0.0, deltat = 1.0;
    unsigned int demTimeStep = 0;
    bool newNeighborList = false;
    particleList particles;
    objectList objects;
    bodyList bodies;
    wallList walls;
    // stress diagnostics of the last DEM step (DEM.h:82-86; filled by refresh when they are on)
    tMat totalStress, contactStress, lubStress, electroStress, hydroStress;
    Context& context() const { return c_; }

private:
    Context& c_;
};

class IO {
public:
    explicit IO(Context& c) : c_(c) {}
    void initialize();                             // IO.cpp:9-112 (the result files of the path)
    void outputStep(LB& lb, DEM& dem);             // IO.cpp:113-168 (screen interval, then createFiles)
    // whether this step's createFiles writes a fluid file that needs node->u (IO.cpp:203-210)
    bool fluidVelocityDue() const;

    std::string workDirectory = "data", configFileName = "config.cfg";
    double realTime = 0.0, maxTime = 0.0, screenExpTime = 0.0;
    double fluidRecycleExpTime = 0.0;              // main.cpp:157
    double fluidExpTime = 0.0, fluidLagrangianExpTime = 0.0;  // main.cpp:148-151
    // main.cpp:155-162: every DEM file's interval is partExpTime
    double partExpTime = 0.0;
    double partRecycleExpTime = 0.0;               // main.cpp:159
    unsigned int currentTimeStep = 0;
    int demSolve = 0, lbmSolve = 0;

private:
    void exportObjectPosition(const objectList& objects, const LB& lb);   // IO.cpp:1316-1338
    void exportObjectForce(const objectList& objects, const LB& lb);      // IO.cpp:1239-1293
    void exportWallForce(const wallList& walls);                          // IO.cpp:1295-1314
    void exportBodyPosition(const bodyList& bodies);                      // IO.cpp:1340-1353
    // the DEM VTK files of createFiles that need no stress diagnostics (IO.cpp:212-256)
    void exportParaviewParticleVelocities(const particleList& particles, const std::string& file);  // IO.cpp:651-700
    void exportParaviewParticles(const particleList& particles, const std::string& file);           // IO.cpp:292-649
    void exportParaviewBoundaries(const LB& lb, const std::string& file);                          // IO.cpp:829-912
    void exportParaviewObjects(const objectList& objects, const std::string& file);                // IO.cpp:702-751
    void exportParaviewBodies(const bodyList& bodies, const std::string& file);                    // IO.cpp:753-784
    void exportParaviewSprings(const bodyList& bodies, const objectList& objects, const std::string& file);  // IO.cpp:786-827
    void exportMass(const LB& lb);                                        // step, time, total fluid mass
    void exportStressTensor(const DEM& dem);                              // IO.cpp:1355-1363
    void exportNormalStressContributions(const DEM& dem);                 // IO.cpp:1365-1386
    void exportShearStressContributions(const DEM& dem);                  // IO.cpp:1388-1409
    void exportRecycleParticles(const particleList& particles, const std::string& file);  // IO.cpp:999-1036
    void createFiles(LB& lb, DEM& dem);                                   // IO.cpp:181-289 (the fluid files)
    void exportRecycleFluid(const LB& lb, const std::string& file);       // IO.cpp:973-997
    void exportEulerianParaviewFluid(const LB& lb, const std::string& file);    // IO.cpp:1139-1235
    void exportLagrangianParaviewFluid(const LB& lb, const std::string& file);  // IO.cpp:1038-1137
    // the lattice fields the fluid files read, downloaded once per output step
    struct FluidState {
        int64_t N = 0;
        std::vector<uint8_t> t, p;
        std::vector<double> f, mass, n, u, visc;
    };
    const FluidState& fluidState(const LB& lb, bool populations, bool velocity);
    Context& c_;
    unsigned lastScreenExp = 0, lastFluidRecycleExp = 0, lastFluidExp = 0, lastFluidLagrangianExp = 0;
    unsigned lastVelExp = 0, lastBoundExp = 0, lastObjectExp = 0, lastBodyExp = 0, lastSpringExp = 0;
    unsigned lastPartRecycleExp = 0, lastPartExp = 0;
    std::string partFileFormat, partRecycleFileFormat, stressTensorFileName, normalStressContFileName, shearStressContFileName;
    std::string velFileFormat, boundFileFormat, objectFileFormat, bodyFileFormat, springFileFormat;
    std::string exportFileName, objectPositionFileName, objectForceFileName, wallForceFileName, massFileName;
    std::string bodyPositionFileName;
    std::string partDirectory, fluidDirectory, fluidRecycleFileFormat, fluidFileFormat, fluidLagrangianFileFormat;
    FluidState fs_;
    unsigned fsStep_ = ~0u;
    bool fsF_ = false, fsU_ = false;
};

// parseConfigFile (main.cpp:128-185) and one goCycle (main.cpp:38-67)
void parseConfigFile(IO& io, DEM& dem, LB& lb, ConfigFile& configFile, CommandLine& commandLine);
void goCycle(IO& io, DEM& dem, LB& lb);

}  // namespace lbdem
