#include "facade.hpp"

#include <sys/stat.h>

#include <cmath>
#include <utility>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>

namespace lbdem {

void tVect::print(std::ostream& os) const { os << x << " " << y << " " << z << " "; }
void tMat::print(std::ostream& os) const {
    os << m00 << " " << m11 << " " << m22 << " " << (m01 + m10) / 2.0 << " " << (m12 + m21) / 2.0 << " "
       << (m02 + m20) / 2.0 << " ";
}
void tMat::printNormal(std::ostream& os) const { os << m00 << " " << m11 << " " << m22 << " "; }
void tMat::printShear(std::ostream& os) const {
    os << (m01 + m10) / 2.0 << " " << (m12 + m21) / 2.0 << " " << (m02 + m20) / 2.0 << " ";
}
void tVect::printFixedLine(std::ostream& os) const {
    os << std::setprecision(8) << std::fixed << x << " " << y << " " << z << "\n";
}

// ------------------------------------------------------------------------------------------
// Context
// ------------------------------------------------------------------------------------------
Context::Context(const std::string& libraryPath, int precision, int device) : api_(Api::load(libraryPath)) {
    params.precision = precision;
    params.device = device;
}

Context::~Context() {
    if (ctx_) api_.destroy(ctx_);
}

void Context::check(int rc, const char* what) const {
    if (rc == LBDEM_OK) return;
    std::cout << "ERROR: " << what << " failed (code " << rc << "): "
              << (ctx_ ? api_.last_error(ctx_) : "no context") << std::endl;
    std::exit(1);
}

void Context::create() {
    if (ctx_) return;
    check(api_.create(&params, &ctx_), "lbdem_create");
    const int np = int(pr.size()), no = int(orad.size());
    check(api_.set_particles(ctx_, np, psize.data(), pr.data(), px0.data()), "lbdem_set_particles");
    check(api_.set_objects(ctx_, no, orad.data(), ox0.data(), ox1.data()), "lbdem_set_objects");
    const int nb = int(brad.size());
    if (nb > 0)
        check(api_.set_bodies(ctx_, nb, brad.data(), bx0.data(), bx1.data(), bk.data(), bL0.data()), "lbdem_set_bodies");
}

void Context::initialise() {
    create();
    if (initialised_) return;
    if (stress) check(api_.set_stress(ctx_, 1), "lbdem_set_stress");
    check(api_.init(ctx_), "lbdem_init");
    initialised_ = true;
}

// ------------------------------------------------------------------------------------------
// configuration (main.cpp:128-185)
// ------------------------------------------------------------------------------------------
void parseConfigFile(IO& io, DEM& dem, LB& lb, ConfigFile& cfg, CommandLine& cl) {
    dem.demInitialRepeat = parse_value(cfg, cl, "demInitialRepeat", 0.0);
    lb.lbmInitialRepeat = parse_value(cfg, cl, "lbmInitialRepeat", 0);
    io.maxTime = parse_value(cfg, cl, "maxTime", 0.0);
    io.screenExpTime = parse_value(cfg, cl, "screenExpTime", 0.0);
    io.fluidExpTime = parse_value(cfg, cl, "fluidExpTime", 0.0);
    io.fluidLagrangianExpTime = parse_value(cfg, cl, "fluidLagrangianExpTime", 0.0);
    io.fluidRecycleExpTime = parse_value(cfg, cl, "fluidRecycleExpTime", 0.0);
    io.partExpTime = parse_value(cfg, cl, "partExpTime", 0.0);
    io.partRecycleExpTime = parse_value(cfg, cl, "partRecycleExpTime", 0.0);
    io.demSolve = parse_value(cfg, cl, "demSolve", 0);
    io.lbmSolve = parse_value(cfg, cl, "lbSolve", 0);
    lb.freeSurface = parse_value(cfg, cl, "freeSurfaceSolve", 0);
    lb.forceField = parse_value(cfg, cl, "forceFieldSolve", 0);
    lb.latticeBoltzmannGet(cfg, cl);
    dem.discreteElementGet(cfg, cl);
    lbdem_params& p = lb.context().params;
    p.demInitialRepeat = dem.demInitialRepeat;
    p.maxTime = io.maxTime;
    p.demSolve = io.demSolve;
    // the stress files of the screen interval and the particle-recycle files need the stress
    // diagnostics, which the reference computes on every DEM step
    lb.context().stress = io.demSolve && (io.screenExpTime > 0 || io.partRecycleExpTime > 0 || io.partExpTime > 0);
    p.lbSolve = io.lbmSolve;
    p.freeSurface = lb.freeSurface;
    p.forceField = lb.forceField;
}

// ------------------------------------------------------------------------------------------
// LB
// ------------------------------------------------------------------------------------------
void LB::latticeBoltzmannGet(ConfigFile& cfg, CommandLine& cl) {
    // keys and defaults of LB.cpp:137-336 (no topography)
    lbdem_params& p = c_.params;
    p.unitLength = parse_value(cfg, cl, "unitLength", 1.0);
    p.unitTime = parse_value(cfg, cl, "unitTime", 1.0);
    p.unitDensity = parse_value(cfg, cl, "unitDensity", 1.0);
    unitLength = p.unitLength;
    unitTime = p.unitTime;
    unitDensity = p.unitDensity;
    const char* ax[3] = {"X", "Y", "Z"};
    for (int a = 0; a < 3; ++a) p.lbSize[a] = parse_value(cfg, cl, std::string("lbSize") + ax[a], 0.0);
    lbRestart = to_int(parse_value(cfg, cl, "lbRestart", "false"), 0);
    lbRestartFile = parse_value(cfg, cl, "lbRestartFile", "");
    if (lbRestart) std::cout << "Computation from LB restart file " << lbRestartFile << std::endl;
    for (int a = 0; a < 3; ++a) p.initVelocity[a] = parse_value(cfg, cl, std::string("initVelocity") + ax[a], 0.0);
    for (int a = 0; a < 3; ++a) p.lbF[a] = parse_value(cfg, cl, std::string("lbF") + ax[a], 0.0);
    for (int b = 0; b < 6; ++b) p.boundary[b] = parse_value(cfg, cl, "boundary" + std::to_string(b), 1);
    p.slipCoefficient = parse_value(cfg, cl, "slipCoefficient", 0.0);
    // rheology (LB.cpp:197-245): an unknown name keeps NEWTONIAN; BINGHAM reads no initVisc and
    // keeps the default (1.0 - 0.5) / 3.0, scaled at LB.cpp:320 like a physical viscosity
    const std::string rheology = parse_value(cfg, cl, "rheologyModel", "none");
    p.rheologyModel = rheology == "BINGHAM" ? LBDEM_RHEO_BINGHAM
                    : rheology == "FRICTIONAL" ? LBDEM_RHEO_FRICTIONAL
                    : rheology == "VOELLMY" ? LBDEM_RHEO_VOELLMY : LBDEM_RHEO_NEWTONIAN;
    p.initVisc = p.rheologyModel == LBDEM_RHEO_BINGHAM ? (1.0 - 0.5) / 3.0 : parse_value(cfg, cl, "initVisc", 1.0);
    if (p.rheologyModel == LBDEM_RHEO_BINGHAM) {
        p.plasticVisc = parse_value(cfg, cl, "plasticVisc", 0.0);
        p.yieldStress = parse_value(cfg, cl, "yieldStress", 0.0);
    }
    if (p.rheologyModel == LBDEM_RHEO_FRICTIONAL || p.rheologyModel == LBDEM_RHEO_VOELLMY)
        p.frictionCoefFluid = parse_value(cfg, cl, "frictionCoefFluid", 0.0);
    if (p.rheologyModel == LBDEM_RHEO_VOELLMY) p.voellmyConstant = parse_value(cfg, cl, "voellmyConstant", 1.0);
    p.turbulenceSolve = to_int(parse_value(cfg, cl, "turbulenceSolve", "0"), 0);
    p.turbConst = parse_value(cfg, cl, "turbConst", 0.0);
    p.fluidHeight = parse_value(cfg, cl, "fluidHeight", 0.0);
    fluidHeight = p.fluidHeight;
}

void LB::latticeDefinition() {
    for (int a = 0; a < 3; ++a) lbSize[a] = unsigned(c_.params.lbSize[a] / c_.params.unitLength);
}

void LB::LBShow() const {
    std::cout << "USED LB PARAMETERS" << std::endl
              << "lbSize: " << lbSize[0] << " " << lbSize[1] << " " << lbSize[2] << std::endl
              << "unitLength " << unitLength << ", unitTime " << unitTime << ", initVisc " << c_.params.initVisc
              << ", fluidHeight " << fluidHeight << ", freeSurface " << freeSurface << ", forceField " << forceField
              << std::endl;
}

void LB::latticeBolzmannInit(wallList&, particleList&, objectList&) {
    if (lbRestart && !c_.initialised()) {
        // the fluidRecycle file (IO.cpp:973-997): lattice size, every node's type, then one
        // record per active node (LB.cpp:352-367, 994-1009, 1098-1138 read it back)
        std::ifstream in(lbRestartFile);
        if (!in.is_open()) {
            std::cout << "ERROR: Condition fluidFileID.is_open() not fulfilled!" << std::endl;
            std::exit(1);
        }
        unsigned sz[3] = {0, 0, 0};
        in >> sz[0] >> sz[1] >> sz[2];
        const size_t nodes = size_t(sz[0]) * sz[1] * sz[2];
        std::vector<uint8_t> types(nodes, 0);
        for (size_t i = 0; i < nodes; ++i) {
            unsigned v = 0;
            in >> v;
            types[i] = uint8_t(v);
        }
        std::vector<double> rec;
        double v = 0.0;
        while (in >> v) rec.push_back(v);
        const int size[3] = {int(sz[0]), int(sz[1]), int(sz[2])};
        c_.create();
        c_.check(c_.api().set_restart_fluid(c_.ctx(), size, types.data(), int(rec.size() / LBDEM_RESTART_REC),
                                            rec.data()),
                 "lbdem_set_restart_fluid");
    }
    c_.initialise();
    lbdem_info in;
    c_.check(c_.api().get_info(c_.ctx(), &in), "lbdem_get_info");
    totalMass = in.totalMassInit;
}

void LB::latticeBolzmannStep(particleList&, wallList&, objectList&) {
    c_.initialise();
    c_.check(c_.api().lb_step(c_.ctx()), "lbdem_lb_step");
}

void LB::latticeBoltzmannCouplingStep(bool& newNeighborList, particleList&, objectList&) {
    c_.initialise();
    c_.check(c_.api().coupling_step(c_.ctx()), "lbdem_coupling_step");
    newNeighborList = false;  // LB.cpp:445
}

void LB::latticeBoltzmannFreeSurfaceStep() {
    c_.initialise();
    c_.check(c_.api().fs_step(c_.ctx()), "lbdem_fs_step");
}

void LB::storeVelocity() {
    c_.initialise();
    c_.check(c_.api().store_velocity(c_.ctx()), "lbdem_store_velocity");
}

double LB::currentMass() const {
    double m = 0.0;
    c_.check(c_.api().total_mass(c_.ctx(), &m), "lbdem_total_mass");
    return m;
}

// ------------------------------------------------------------------------------------------
// DEM
// ------------------------------------------------------------------------------------------
static std::ifstream open_input(const char* path) {
    std::ifstream in(path);
    std::cout << "Reading " << path << std::endl;
    if (!in.is_open()) {
        std::cout << "ERROR: Condition " << path << " is_open() not fulfilled!" << std::endl;
        std::exit(1);
    }
    return in;
}

void DEM::discreteElementGet(ConfigFile& cfg, CommandLine& cl) {
    // keys and defaults of DEM.cpp:19-75
    lbdem_params& p = c_.params;
    p.density = parse_value(cfg, cl, "density", 0.0);
    p.objectDensity = parse_value(cfg, cl, "objectDensity", 0.0);
    p.bodyDensity = parse_value(cfg, cl, "bodyDensity", 0.0);
    p.linearStiff = parse_value(cfg, cl, "linearStiff", 0.0);
    p.frictionCoefPart = parse_value(cfg, cl, "frictionCoefPart", 0.0);
    p.lubNormal = parse_value(cfg, cl, "lubNormal", 0.0);
    p.lubTangential = parse_value(cfg, cl, "lubTangential", 0.0);
    p.lubRotation = parse_value(cfg, cl, "lubRotation", 0.0);
    p.lubCutoff = parse_value(cfg, cl, "lubCutoff", 0.0);
    p.kinVisc = parse_value(cfg, cl, "kinVisc", 0.0);
    p.electroForce = parse_value(cfg, cl, "electroForce", 0.0);
    p.debyeLength = parse_value(cfg, cl, "debyeLength", 0.0);
    p.objWaitTime = parse_value(cfg, cl, "objWaitTime", 0);
    p.springDamp = parse_value(cfg, cl, "springDamp", 0.0);
    p.epsilon = parse_value(cfg, cl, "epsilon", 0.0);
    p.nonLinType = parse_value(cfg, cl, "nonLinType", 0);
    p.k13 = parse_value(cfg, cl, "k13", 0.0);
    const double scale = parse_value(cfg, cl, "scale", 1.0);

    // input/particles.dat: count, then index size r x y z vx vy vz wx wy wz q0(4) q1(4) (DEM.cpp:81-134);
    // the velocity columns are read and dropped: particle::initialize resets x1 and w1
    // (elmt.cpp:28-31, called from DEM.cpp:232)
    {
        std::ifstream in = open_input("./input/particles.dat");
        unsigned n = 0;
        in >> n;
        c_.psize.assign(n, 1);
        c_.pr.assign(n, 0.0);
        c_.px0.assign(3 * size_t(n), 0.0);
        for (unsigned i = 0; i < n; ++i) {
            double index, size, r, x[3], v[3], w[3], q[8];
            in >> index >> size >> r >> x[0] >> x[1] >> x[2] >> v[0] >> v[1] >> v[2] >> w[0] >> w[1] >> w[2];
            for (double& t : q) in >> t;
            c_.psize[i] = int(size);
            c_.pr[i] = r * scale;
            for (int a = 0; a < 3; ++a) {
                c_.px0[3 * i + a] = x[a] * scale - 0.05;
            }
        }
    }
    // input/objects.dat: count, then index r x y z vx vy vz (DEM.cpp:136-165)
    {
        std::ifstream in = open_input("./input/objects.dat");
        unsigned n = 0;
        in >> n;
        c_.orad.assign(n, 0.0);
        c_.ox0.assign(3 * size_t(n), 0.0);
        c_.ox1.assign(3 * size_t(n), 0.0);
        for (unsigned i = 0; i < n; ++i) {
            double index, r, x[3], v[3];
            in >> index >> r >> x[0] >> x[1] >> x[2] >> v[0] >> v[1] >> v[2];
            c_.orad[i] = r * scale;
            for (int a = 0; a < 3; ++a) {
                c_.ox0[3 * i + a] = x[a] * scale;
                c_.ox1[3 * i + a] = v[a];
            }
        }
    }
    // input/body.dat: count, then index r x y z vx vy vz k L0 (DEM.cpp:168-196)
    {
        std::ifstream in = open_input("./input/body.dat");
        unsigned n = 0;
        in >> n;
        c_.brad.assign(n, 0.0); c_.bk.assign(n, 0.0); c_.bL0.assign(n, 0.0);
        c_.bx0.assign(3 * size_t(n), 0.0);
        c_.bx1.assign(3 * size_t(n), 0.0);
        for (unsigned i = 0; i < n; ++i) {
            double index, r, x[3], v[3], k, L0;
            in >> index >> r >> x[0] >> x[1] >> x[2] >> v[0] >> v[1] >> v[2] >> k >> L0;
            c_.brad[i] = r;
            c_.bk[i] = k;
            c_.bL0[i] = L0;
            for (int a = 0; a < 3; ++a) {
                c_.bx0[3 * i + a] = x[a] * scale;
                c_.bx1[3 * i + a] = v[a];
            }
        }
    }
}

void DEM::discreteElementInit(const unsIntList&, const doubleList&, const vecList&, const double& externalTimeStep,
                              const tVect, const double&, bool&) {
    deltat = externalTimeStep / double(1);  // one DEM step per LB step (DEM.cpp:327)
    c_.create();
    newNeighborList = true;
}

void DEM::discreteElementStep(const unsIntList&) {
    c_.initialise();
    c_.check(c_.api().dem_step(c_.ctx()), "lbdem_dem_step");
    ++demTimeStep;        // DEM.cpp:365-366
    demTime += deltat;
}

static tVect rec3(const double* r, int off) { return tVect(r[off], r[off + 1], r[off + 2]); }

void DEM::refresh() {
    if (!c_.initialised()) return;
    lbdem_info in;
    c_.check(c_.api().get_info(c_.ctx(), &in), "lbdem_get_info");
    std::vector<double> rec;
    rec.resize(size_t(std::max(in.nParticles, 1)) * LBDEM_P_REC);
    c_.check(c_.api().get_particles(c_.ctx(), rec.data()), "lbdem_get_particles");
    particles.resize(in.nParticles);
    for (int i = 0; i < in.nParticles; ++i) {
        const double* r = &rec[size_t(i) * LBDEM_P_REC];
        particle& P = particles[i];
        P.index = unsigned(i);
        P.x0 = rec3(r, LBDEM_P_X0); P.x1 = rec3(r, LBDEM_P_X1); P.w1 = rec3(r, LBDEM_P_W1);
        P.FHydro = rec3(r, LBDEM_P_FHYDRO); P.MHydro = rec3(r, LBDEM_P_MHYDRO);
        P.FContact = rec3(r, LBDEM_P_FCONTACT); P.MContact = rec3(r, LBDEM_P_MCONTACT);
        P.FLub = rec3(r, LBDEM_P_FLUB); P.MLub = rec3(r, LBDEM_P_MLUB); P.FRep = rec3(r, LBDEM_P_FREP);
        P.FTotal = rec3(r, LBDEM_P_FTOTAL); P.MTotal = rec3(r, LBDEM_P_MTOTAL); P.FGrav = rec3(r, LBDEM_P_FGRAV);
        P.fluidVolume = r[LBDEM_P_FLUIDVOL]; P.xiw = rec3(r, LBDEM_P_XIW); P.xio = rec3(r, LBDEM_P_XIO);
        P.r = r[LBDEM_P_R]; P.m = r[LBDEM_P_M];
    }
    if (c_.stress) {  // DEM.cpp:745-771, 2705
        std::vector<double> g(45), st(size_t(std::max(in.nParticles, 1)) * 9), dp(size_t(std::max(in.nParticles, 1)) * 3);
        c_.check(c_.api().get_stress(c_.ctx(), g.data(), st.data(), dp.data()), "lbdem_get_stress");
        tMat* t[5] = {&totalStress, &hydroStress, &lubStress, &contactStress, &electroStress};
        for (int k = 0; k < 5; ++k) t[k]->set(&g[9 * size_t(k)]);
        std::vector<double> ps(size_t(std::max(in.nParticles, 1)) * LBDEM_S_REC);
        c_.check(c_.api().get_particle_stress(c_.ctx(), ps.data()), "lbdem_get_particle_stress");
        for (int i = 0; i < in.nParticles; ++i) {
            particle& P = particles[i];
            const double* R = &ps[size_t(i) * LBDEM_S_REC];
            P.STotal.set(R + LBDEM_S_STOTAL); P.SElas.set(R + LBDEM_S_SELAS);
            P.SVisc.set(R + LBDEM_S_SVISC); P.SVisc_h.set(R + LBDEM_S_SVISC_H);
            P.overlap = rec3(R, LBDEM_S_OVERLAP);
            P.disp = rec3(R, LBDEM_S_DISP);
            P.coord = R[LBDEM_S_COORD]; P.lub_coord = R[LBDEM_S_LUB_COORD];
            // DEM.cpp:751, computed by every DEM step (0 until the first)
            P.pressure = in.demTimeStep > 0 ? -1.0 / 3.0 * (P.STotal.m00 + P.STotal.m11 + P.STotal.m22) : 0.0;
            P.disp_r = std::sqrt(P.disp.x * P.disp.x + P.disp.z * P.disp.z);       // DEM.cpp:2712
        }
    }
    rec.assign(size_t(std::max(in.nObjects, 1)) * LBDEM_O_REC, 0.0);
    c_.check(c_.api().get_objects(c_.ctx(), rec.data()), "lbdem_get_objects");
    objects.resize(in.nObjects);
    for (int i = 0; i < in.nObjects; ++i) {
        const double* r = &rec[size_t(i) * LBDEM_O_REC];
        object& O = objects[i];
        O.index = unsigned(i);
        O.x0 = rec3(r, LBDEM_O_X0); O.x1 = rec3(r, LBDEM_O_X1); O.w1 = rec3(r, LBDEM_O_W1);
        O.FHydro = rec3(r, LBDEM_O_FHYDRO); O.MHydro = rec3(r, LBDEM_O_MHYDRO);
        O.FContact = rec3(r, LBDEM_O_FCONTACT); O.MContact = rec3(r, LBDEM_O_MCONTACT);
        O.FLub = rec3(r, LBDEM_O_FLUB); O.MLub = rec3(r, LBDEM_O_MLUB);
        O.FTotal = rec3(r, LBDEM_O_FTOTAL); O.MTotal = rec3(r, LBDEM_O_MTOTAL); O.FGrav = rec3(r, LBDEM_O_FGRAV);
        O.FRep = tVect();  // no electrostatic force acts on objects (DEM.cpp:786-790)
        O.r = r[LBDEM_O_R]; O.m = r[LBDEM_O_M];
    }
    rec.assign(size_t(std::max(in.nBodies, 1)) * LBDEM_B_REC, 0.0);
    c_.check(c_.api().get_bodies(c_.ctx(), rec.data()), "lbdem_get_bodies");
    bodies.resize(in.nBodies);
    for (int i = 0; i < in.nBodies; ++i) {
        const double* r = &rec[size_t(i) * LBDEM_B_REC];
        body& B = bodies[i];
        B.index = unsigned(i);
        B.x0 = rec3(r, LBDEM_B_X0); B.x1 = rec3(r, LBDEM_B_X1); B.FSpring = rec3(r, LBDEM_B_FSPRING);
        B.FTotal = rec3(r, LBDEM_B_FTOTAL); B.FGrav = rec3(r, LBDEM_B_FGRAV);
        B.L = r[LBDEM_B_L]; B.L0 = r[LBDEM_B_L0]; B.k = r[LBDEM_B_K]; B.r = r[LBDEM_B_R]; B.m = r[LBDEM_B_M];
    }
    rec.assign(size_t(std::max(in.nWalls, 1)) * LBDEM_W_REC, 0.0);
    c_.check(c_.api().get_walls(c_.ctx(), rec.data()), "lbdem_get_walls");
    walls.resize(in.nWalls);
    for (int i = 0; i < in.nWalls; ++i) {
        const double* r = &rec[size_t(i) * LBDEM_W_REC];
        wall& W = walls[i];
        W.index = unsigned(i);
        W.FHydro = rec3(r, LBDEM_W_FHYDRO); W.FContact = rec3(r, LBDEM_W_FCONTACT);
        W.FLub = rec3(r, LBDEM_W_FLUB); W.FTotal = rec3(r, LBDEM_W_FTOTAL);
    }
}

// ------------------------------------------------------------------------------------------
// IO: the result files of the path (headers and columns as IO.cpp:50-85, 1239-1338) plus mass.dat
// ------------------------------------------------------------------------------------------
static void append_line(const std::string& path, const char* text) {
    std::ofstream f(path, std::ios::app);
    f << text;
}

void IO::initialize() {
    partDirectory = workDirectory + "/particleData";
    fluidDirectory = workDirectory + "/fluidData";
    if (demSolve) {
        const int a = ::mkdir(partDirectory.c_str(), 0777);
        std::cout << "Work directory created = " << partDirectory << ". Result: " << a << "\n";
    }
    if (lbmSolve) {
        const int a = ::mkdir(fluidDirectory.c_str(), 0777);
        std::cout << "Work directory created = " << fluidDirectory << ". Result: " << a << "\n";
    }
    fluidFileFormat = fluidDirectory + "/fluid%010u.vti";
    fluidLagrangianFileFormat = fluidDirectory + "/fluidLag%010u.vtu";
    fluidRecycleFileFormat = fluidDirectory + "/fluidRecycle%010u.dat";
    velFileFormat = partDirectory + "/vel%04u.vtk";
    boundFileFormat = partDirectory + "/bound%04u.vtk";
    objectFileFormat = partDirectory + "/object%04u.vtk";
    bodyFileFormat = partDirectory + "/body%04u.vtk";
    springFileFormat = partDirectory + "/spring%04u.vtk";
    partRecycleFileFormat = partDirectory + "/partRecycle%04u.dat";
    partFileFormat = partDirectory + "/part%04u.vtk";
    exportFileName = workDirectory + "/export.dat";
    if (demSolve) {
        objectPositionFileName = workDirectory + "/objPosition.dat";
        append_line(objectPositionFileName,
                    "position(x) position(y) position(z) velocity(x) velocity(y) velocity(z) depth(y)  \n");
        bodyPositionFileName = workDirectory + "/bodyPosition.dat";
        append_line(bodyPositionFileName,
                    "position(x) position(y) position(z) velocity(x) velocity(y) velocity(z) Length Delta  \n");
        objectForceFileName = workDirectory + "/objForce.dat";
        append_line(objectForceFileName,
                    "FExerted(x) FExerted(y) FExerted(z) SExerted(x) SExerted(y) SExerted(z) FTotal(x) FTotal(y) "
                    "FTotal(z) FContact(x) FContact(y) FContact(z) FHydro(x) FHydro(y) FHydro(z) FLub(x) FLub(y) "
                    "FLub(z) FRep(x) FRep(y) FRep(z) FGrav(x) FGrav(y) FGrav(z) \n");
        wallForceFileName = workDirectory + "/wallForce.dat";
        append_line(wallForceFileName,
                    "0-FTotal(x) 0-FTotal(y) 0-FTotal(z) 1-FTotal(x) 1-FTotal(y) 1-FTotal(z) 2-FTotal(x) "
                    "2-FTotal(y) 2-FTotal(z) 3-FTotal(x) 3-FTotal(y) 3-FTotal(z) 4-FTotal(x) 4-FTotal(y) "
                    "4-FTotal(z) 5-FTotal(x) 5-FTotal(y) 5-FTotal(z) \n");
        stressTensorFileName = workDirectory + "/stressTensor.dat";
        append_line(stressTensorFileName, "xx yy zz xy yz zx \n");
        normalStressContFileName = workDirectory + "/normalStressCont.dat";
        append_line(normalStressContFileName,
                    "hydro-xx hydro-yy hydro-zz lub-xx lub-yy lub-zz contact-xx contact-yy contact-zz rep-xx rep-yy "
                    "rep-zz \n");
        shearStressContFileName = workDirectory + "/shearStressCont.dat";
        append_line(shearStressContFileName,
                    "hydro-xy hydro-yz hydro-zx lub-xy lub-yz lub-zx contact-xy contact-yz contact-zx rep-xy rep-yz "
                    "rep-zx \n");
    }
    if (lbmSolve) {
        massFileName = workDirectory + "/mass.dat";
        append_line(massFileName, "step time totalFluidMass\n");
    }
    lastScreenExp = 0;
    lastFluidRecycleExp = lastFluidExp = lastFluidLagrangianExp = 0;
    lastVelExp = lastBoundExp = lastObjectExp = lastBodyExp = lastSpringExp = 0;
    lastPartRecycleExp = lastPartExp = 0;
}

void IO::outputStep(LB& lb, DEM& dem) {
    const unsigned screenExpCounter = screenExpTime > 0 ? unsigned(realTime / screenExpTime) + 1 : 0;
    if (screenExpCounter > lastScreenExp) {
        lastScreenExp = screenExpCounter;
        std::ofstream exportFile(exportFileName, std::ios::app);
        std::cout << currentTimeStep << "; time=" << realTime << "\t";
        exportFile << currentTimeStep << "; time=" << realTime << "\t";
        dem.refresh();
        if (demSolve) {
            if (!dem.objects.empty()) {
                exportObjectPosition(dem.objects, lb);
                exportObjectForce(dem.objects, lb);
            }
            if (!dem.bodies.empty()) exportBodyPosition(dem.bodies);
            if (!dem.walls.empty()) exportWallForce(dem.walls);
            exportStressTensor(dem);
            exportNormalStressContributions(dem);
            exportShearStressContributions(dem);
        }
        if (lbmSolve && c_.initialised()) exportMass(lb);
        std::cout << "\n";
        exportFile << "\n";
        std::cout.flush();
    }
    createFiles(lb, dem);
}

// an output interval's counter (IO.cpp:117, 186-207)
static unsigned exp_counter(double realTime, double expTime) {
    return expTime > 0 ? unsigned(realTime / expTime) + 1 : 0;
}

bool IO::fluidVelocityDue() const {
    return lbmSolve && (exp_counter(realTime, fluidLagrangianExpTime) > lastFluidLagrangianExp ||
                        exp_counter(realTime, fluidRecycleExpTime) > lastFluidRecycleExp);
}

void IO::createFiles(LB& lb, DEM& dem) {
    if (!c_.initialised()) return;
    char path[1024];
    if (demSolve) {
        // IO.cpp:212-283 (one interval, partExpTime, for the VTK files; partRecycleExpTime for the
        // particle-recycle file, numbered by its counter); the particle and force-chain files are not
        // written
        const unsigned cv = exp_counter(realTime, partExpTime), cr = exp_counter(realTime, partRecycleExpTime);
        const bool due = cv > lastVelExp || cv > lastPartExp || cv > lastBoundExp ||
                         (!dem.objects.empty() && cv > lastObjectExp) || cv > lastBodyExp || cv > lastSpringExp ||
                         cr > lastPartRecycleExp;
        if (due) dem.refresh();
        if (cv > lastVelExp) {
            lastVelExp = cv;
            std::snprintf(path, sizeof(path), velFileFormat.c_str(), currentTimeStep);
            exportParaviewParticleVelocities(dem.particles, path);
        }
        if (cv > lastPartExp) {
            lastPartExp = cv;
            std::snprintf(path, sizeof(path), partFileFormat.c_str(), currentTimeStep);
            exportParaviewParticles(dem.particles, path);
        }
        if (cv > lastBoundExp) {
            lastBoundExp = cv;
            std::snprintf(path, sizeof(path), boundFileFormat.c_str(), currentTimeStep);
            exportParaviewBoundaries(lb, path);
        }
        if (!dem.objects.empty() && cv > lastObjectExp) {
            lastObjectExp = cv;
            std::snprintf(path, sizeof(path), objectFileFormat.c_str(), currentTimeStep);
            exportParaviewObjects(dem.objects, path);
        }
        if (cv > lastBodyExp) {
            lastBodyExp = cv;
            std::snprintf(path, sizeof(path), bodyFileFormat.c_str(), currentTimeStep);
            exportParaviewBodies(dem.bodies, path);
        }
        if (cv > lastSpringExp) {
            lastSpringExp = cv;
            std::snprintf(path, sizeof(path), springFileFormat.c_str(), currentTimeStep);
            exportParaviewSprings(dem.bodies, dem.objects, path);
        }
        if (cr > lastPartRecycleExp) {
            lastPartRecycleExp = cr;
            std::snprintf(path, sizeof(path), partRecycleFileFormat.c_str(), cr);
            exportRecycleParticles(dem.particles, path);
        }
    }
    if (!lbmSolve) return;
    unsigned counter = exp_counter(realTime, fluidExpTime);
    if (counter > lastFluidExp) {
        lastFluidExp = counter;
        std::snprintf(path, sizeof(path), fluidFileFormat.c_str(), currentTimeStep);
        exportEulerianParaviewFluid(lb, path);
    }
    counter = exp_counter(realTime, fluidLagrangianExpTime);
    if (counter > lastFluidLagrangianExp) {
        lastFluidLagrangianExp = counter;
        std::snprintf(path, sizeof(path), fluidLagrangianFileFormat.c_str(), currentTimeStep);
        exportLagrangianParaviewFluid(lb, path);
    }
    counter = exp_counter(realTime, fluidRecycleExpTime);
    if (counter > lastFluidRecycleExp) {
        lastFluidRecycleExp = counter;
        std::snprintf(path, sizeof(path), fluidRecycleFileFormat.c_str(), currentTimeStep);
        exportRecycleFluid(lb, path);
    }
}

const IO::FluidState& IO::fluidState(const LB& lb, bool populations, bool velocity) {
    if (fsStep_ != currentTimeStep) {
        fsStep_ = currentTimeStep;
        fsF_ = fsU_ = false;
        Context& c = lb.context();
        lbdem_info in;
        c.check(c.api().get_info(c.ctx(), &in), "lbdem_get_info");
        fs_.N = in.totNodes;
        fs_.t.assign(fs_.N, 0);
        fs_.p.assign(fs_.N, 0);
        fs_.mass.assign(fs_.N, 0.0);
        fs_.n.assign(fs_.N, 0.0);
        c.check(c.api().get_types(c.ctx(), fs_.t.data(), fs_.p.data(), nullptr), "lbdem_get_types");
        c.check(c.api().get_fields(c.ctx(), nullptr, fs_.mass.data(), fs_.n.data()), "lbdem_get_fields");
    }
    Context& c = lb.context();
    if (populations && !fsF_) {
        fs_.f.assign(size_t(Q_DIRECTIONS) * fs_.N, 0.0);
        c.check(c.api().get_fields(c.ctx(), fs_.f.data(), nullptr, nullptr), "lbdem_get_fields");
        fsF_ = true;
    }
    if (velocity && !fsU_) {
        fs_.u.assign(3 * size_t(fs_.N), 0.0);
        fs_.visc.assign(fs_.N, 0.0);
        c.check(c.api().get_velocity(c.ctx(), fs_.u.data(), fs_.visc.data()), "lbdem_get_velocity");
        fsU_ = true;
    }
    return fs_;
}

// every node's type (1 inside a particle) and physical mass (0 where no node holds fluid), as
// IO.cpp:1139-1235 writes them, the closing tags straight after the last value
void IO::exportEulerianParaviewFluid(const LB& lb, const std::string& file) {
    const FluidState& st = fluidState(lb, false, false);
    const double zero = 0.0;
    std::ofstream out(file);
    out << "<VTKFile type=\"ImageData\" version=\"0.1\">\n";
    out << " <ImageData WholeExtent=\"0 " << lb.lbSize[0] - 1 << " 0 " << lb.lbSize[1] - 1 << " 0 " << lb.lbSize[2] - 1
        << "\" " << "Origin=\"0.0 0.0 0.0\" Spacing=\"" << lb.unitLength << " " << lb.unitLength << " " << lb.unitLength
        << "\">\n";
    out << "  <Piece Extent=\"0 " << lb.lbSize[0] - 1 << " 0 " << lb.lbSize[1] - 1 << " 0 " << lb.lbSize[2] - 1 << "\">\n";
    out << "   <PointData>\n";
    out << "    <DataArray type=\"Int8\" Name=\"type\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"8\">\n";
    for (int64_t i = 0; i < st.N; ++i) out << (st.p[i] ? 1u : unsigned(st.t[i])) << " ";
    out << "    </DataArray>\n";
    out << "    <DataArray type=\"Float64\" Name=\"amass\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"2\">\n";
    for (int64_t i = 0; i < st.N; ++i) {
        if (st.t[i] == 0 || st.t[i] == 3) out << st.mass[i] * lb.unitDensity << " ";
        else out << zero << " ";
    }
    out << "    </DataArray>\n";
    out << "   </PointData>\n";
    out << "   <CellData>\n";
    out << "   </CellData>\n";
    out << "  </Piece>\n";
    out << " </ImageData>\n";
    out << "</VTKFile>\n";
}

// the active nodes (ascending index) as points: velocity, pressure, density, viscosity, mass, type
// and position in physical units, as IO.cpp:1038-1137 writes them
void IO::exportLagrangianParaviewFluid(const LB& lb, const std::string& file) {
    const FluidState& st = fluidState(lb, false, true);
    std::vector<int64_t> act;
    for (int64_t i = 0; i < st.N; ++i)
        if (st.t[i] == 0 || st.t[i] == 3) act.push_back(i);
    const double speed = lb.unitSpeed(), pressure = lb.unitPressure(), initDensity = 1.0;
    const unsigned X = lb.lbSize[0], Y = lb.lbSize[1];
    std::ofstream out(file);
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
           "compressor=\"vtkZLibDataCompressor\">\n";
    out << " <UnstructuredGrid GhostLevel=\"0\">\n";
    out << "  <Piece NumberOfPoints=\"" << act.size() << "\" NumberOfCells=\"" << act.size() << "\">\n";
    out << "   <PointData>\n";
    out << "    <DataArray type=\"Float64\" Name=\"v\" NumberOfComponents=\"3\" format=\"ascii\"/>\n";
    for (int64_t i : act) out << st.u[3 * i] * speed << " " << st.u[3 * i + 1] * speed << " " << st.u[3 * i + 2] * speed << "\n";
    out << "    <DataArray type=\"Float64\" Name=\"pressure\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"2\"/>\n";
    for (int64_t i : act) out << 0.3333333 * (st.n[i] - initDensity) * pressure << "\n";
    out << "    <DataArray type=\"Float64\" Name=\"n\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"2\"/>\n";
    for (int64_t i : act) out << st.n[i] << "\n";
    out << "    <DataArray type=\"Float64\" Name=\"visc\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"2\"/>\n";
    for (int64_t i : act) out << st.visc[i] << "\n";
    out << "    <DataArray type=\"Float64\" Name=\"mass\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"2\"/>\n";
    for (int64_t i : act) out << st.mass[i] * lb.unitDensity << "\n";
    out << "    <DataArray type=\"Int8\" Name=\"type\" format=\"ascii\" RangeMin=\"0\" RangeMax=\"8\"/>\n";
    for (int64_t i : act) out << unsigned(st.t[i]) << "\n";
    out << "   </PointData>\n";
    out << "   <CellData>\n";
    out << "   </CellData>\n";
    out << "   <Points>\n";
    out << "    <DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\"/>\n";
    for (int64_t i : act) {
        const double x = double(unsigned(i % X)), y = double(unsigned((i / X) % Y)), z = double(unsigned(i / (int64_t(X) * Y)));
        out << x * lb.unitLength << " " << y * lb.unitLength << " " << z * lb.unitLength << "\n";
    }
    out << "   </Points>\n";
    out << "   <Cells>\n";
    out << "    <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
    for (size_t i = 1; i < act.size() + 1; ++i) out << i - 1 << "\n";
    out << "    </DataArray>\n";
    out << "    <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
    for (size_t i = 1; i < act.size() + 1; ++i) out << i << "\n";
    out << "    </DataArray>\n";
    out << "    <DataArray type=\"Int64\" Name=\"types\" format=\"ascii\">\n";
    for (size_t i = 0; i < act.size(); ++i) out << 1 << "\n";
    out << "    </DataArray>\n";
    out << "   </Cells>\n";
    out << "  </Piece>\n";
    out << " </UnstructuredGrid>\n";
    out << "</VTKFile>";
}

// sizes, every node's type, then per active node (ascending index, as LB::cleanLists sorts
// activeNodes): n - 1, u.x, u.y, u.z (u.dot of the unit vectors), mass - 1, visc, f[0..18], each
// followed by a space, with the stream's default formatting (initDensity is 1, LB.cpp:330)
void IO::exportRecycleFluid(const LB& lb, const std::string& file) {
    const FluidState& st = fluidState(lb, true, true);
    const int64_t N = st.N;
    const std::vector<uint8_t>& t = st.t;
    const std::vector<double>&f = st.f, &mass = st.mass, &n = st.n, &u = st.u, &visc = st.visc;
    std::ofstream out(file);
    out << lb.lbSize[0] << " " << lb.lbSize[1] << " " << lb.lbSize[2] << "\n";
    for (int64_t i = 0; i < N; ++i) out << unsigned(t[i]) << " ";
    out << std::endl;
    std::cout << std::endl;
    const double initDensity = 1.0;
    const tVect Xp(1.0, 0.0, 0.0), Yp(0.0, 1.0, 0.0), Zp(0.0, 0.0, 1.0);
    for (int64_t i = 0; i < N; ++i) {
        if (t[i] != 0 && t[i] != 3) continue;
        const tVect ui(u[3 * i], u[3 * i + 1], u[3 * i + 2]);
        out << n[i] - initDensity << " " << ui.dot(Xp) << " " << ui.dot(Yp) << " " << ui.dot(Zp) << " "
            << mass[i] - initDensity << " " << visc[i] << " ";
        for (int j = 0; j < Q_DIRECTIONS; ++j) out << f[size_t(j) * N + i] << " ";
        out << std::endl;
    }
}

void IO::exportObjectPosition(const objectList& objects, const LB& lb) {
    std::ofstream f(objectPositionFileName, std::ios::app);
    for (const object& o : objects) {
        const double depth = (lb.fluidHeight + o.r) - o.x0.y;
        o.x0.print(f);
        f << "\t";
        o.x1.print(f);
        f << "\t" << depth;
        f << std::endl;
    }
}

// legacy VTK polydata of the particle centres with their velocities
void IO::exportParaviewParticleVelocities(const particleList& particles, const std::string& file) {
    std::ofstream out(file);
    const int n = int(particles.size());
    out << "# vtk DataFile Version 3.0\n" << "Data.vtk\n" << "ASCII\n" << "DATASET POLYDATA\n";
    out << "POINTS " << n << " float\n";
    for (const particle& p : particles) p.x0.printFixedLine(out);
    out << "\n";
    out << " POINT_DATA " << n << "\n";
    out << "VECTORS velocity float\n";
    for (const particle& p : particles) p.x1.printFixedLine(out);
    out << "\n";
}

// the particles with their diameters, displacements, stresses, overlaps and contact counts, one
// scalar array each in the reference's order and signs (y is the vertical: "z" in the names)
void IO::exportParaviewParticles(const particleList& particles, const std::string& file) {
    std::ofstream out(file);
    const int n = int(particles.size());
    out << "# vtk DataFile Version 3.0\n" << "Data.vtk\n" << "ASCII\n" << "DATASET POLYDATA\n" << "  \n";
    out << "POINTS " << n << " float\n";
    for (const particle& p : particles) p.x0.printFixedLine(out);
    out << "\n";
    out << "POINT_DATA " << n << "\n";
    using F = double (*)(const particle&);
    static const std::pair<const char*, F> scalars[] = {
        {"diameter", [](const particle& p) { return 2.0 * p.r; }},
        {"z-displacement", [](const particle& p) { return -p.disp.y; }},
        {"x-displacement", [](const particle& p) { return -p.disp.x; }},
        {"y-displacement", [](const particle& p) { return -p.disp.z; }},
        {"z-vel", [](const particle& p) { return -p.x1.y; }},
        {"r-displacement", [](const particle& p) { return p.disp_r; }},
        {"sig-xx", [](const particle& p) { return p.STotal.m00; }},
        {"sig-yy", [](const particle& p) { return p.STotal.m22; }},
        {"pressure", [](const particle& p) { return p.pressure; }},
        {"zz-stress", [](const particle& p) { return -p.STotal.m11; }},
        {"xz-stress", [](const particle& p) { return -(p.STotal.m01 + p.STotal.m10) * 0.5; }},
        {"xx-stress-el", [](const particle& p) { return -p.SElas.m00; }},
        {"yy-stress-el", [](const particle& p) { return -p.SElas.m22; }},
        {"zz-stress-el", [](const particle& p) { return -p.SElas.m11; }},
        {"xz-stress-el", [](const particle& p) { return -(p.SElas.m01 + p.SElas.m10) * 0.5; }},
        {"yz-stress-el", [](const particle& p) { return -(p.SElas.m12 + p.SElas.m21) * 0.5; }},
        {"xy-stress-el", [](const particle& p) { return -(p.SElas.m02 + p.SElas.m20) * 0.5; }},
        {"xx-stress-vis", [](const particle& p) { return -p.SVisc.m00; }},
        {"yy-stress-vis", [](const particle& p) { return -p.SVisc.m22; }},
        {"zz-stress-vis", [](const particle& p) { return -p.SVisc.m11; }},
        {"xz-stress-vis", [](const particle& p) { return -(p.SVisc.m01 + p.SVisc.m10) * 0.5; }},
        {"xy-stress-vis", [](const particle& p) { return -(p.SVisc.m02 + p.SVisc.m20) * 0.5; }},
        {"yz-stress-vis", [](const particle& p) { return -(p.SVisc.m12 + p.SVisc.m21) * 0.5; }},
        {"zz-stress-vis-h", [](const particle& p) { return -p.SVisc_h.m11; }},
        {"xz-stress-vis-h", [](const particle& p) { return -(p.SVisc_h.m01 + p.SVisc_h.m10) * 0.5; }},
        {"xy-stress", [](const particle& p) { return -(p.STotal.m02 + p.STotal.m20) * 0.5; }},
        {"yz-stress", [](const particle& p) { return -(p.STotal.m12 + p.STotal.m21) * 0.5; }},
        {"delta-z", [](const particle& p) { return p.overlap.y; }},
        {"delta-x", [](const particle& p) { return p.overlap.x; }},
        {"delta-y", [](const particle& p) { return p.overlap.z; }},
        {"coord", [](const particle& p) { return p.coord; }},
        {"lub_coord", [](const particle& p) { return p.lub_coord; }},
    };
    for (const auto& s : scalars) {
        out << "SCALARS " << s.first << " float\n";
        out << "LOOKUP_TABLE default \n";
        for (const particle& p : particles) out << s.second(p) << "\n";
        out << "\n";
    }
}

// the 12 edges of the DEM box (x, fluid height with a free surface, z) as line cells
void IO::exportParaviewBoundaries(const LB& lb, const std::string& file) {
    std::ofstream out(file);
    const double L = lb.unitLength;
    const double X = double(lb.lbSize[0]) * L, Z = double(lb.lbSize[2]) * L;
    const double Y = lb.freeSurface ? lb.fluidHeight : double(lb.lbSize[1]) * L;
    out << "# vtk DataFile Version 3.0\n" << "DataFC.vtk\n" << "ASCII\n" << "DATASET POLYDATA\n" << "\n";
    out << "POINTS " << 24 << " float\n";
    const double corner[24][3] = {{0.0, 0.0, 0.0}, {X, 0.0, 0.0}, {X, 0.0, 0.0}, {X, 0.0, Z}, {X, 0.0, Z}, {0.0, 0.0, Z},
                                  {0.0, 0.0, Z}, {0.0, 0.0, 0.0},
                                  {0.0, 0.0, 0.0}, {0.0, Y, 0.0}, {X, 0.0, 0.0}, {X, Y, 0.0}, {X, 0.0, Z}, {X, Y, Z},
                                  {0.0, 0.0, Z}, {0.0, Y, Z},
                                  {0.0, Y, 0.0}, {X, Y, 0.0}, {X, Y, 0.0}, {X, Y, Z}, {X, Y, Z}, {0.0, Y, Z},
                                  {0.0, Y, Z}, {0.0, Y, 0.0}};
    for (const auto& c : corner) out << c[0] << " " << c[1] << " " << c[2] << "\n";
    out << "\n";
    out << "LINES " << 12 << " " << 36 << "\n";
    for (int i = 0; i < 12; i++) out << "2 " << 2 * i << " " << 2 * i + 1 << "\n";
    out << "\n";
    out << "POINT_DATA " << 24 << "\n";
    out << "SCALARS radius float\n";
    out << "LOOKUP_TABLE default \n";
    for (int i = 0; i < 24; ++i) out << "1.0\n";
}

// impactor centres as vertex cells with their diameters
void IO::exportParaviewObjects(const objectList& objects, const std::string& file) {
    std::ofstream out(file);
    const int n = int(objects.size());
    out << "# vtk DataFile Version 3.0\n" << "Data.vtk\n" << "ASCII\n" << "DATASET UNSTRUCTURED_GRID\n" << "  \n";
    out << "POINTS " << n << " float\n";
    for (const object& o : objects) o.x0.printFixedLine(out);
    out << "\n";
    out << "CELL_TYPES " << n << "\n";
    for (int i = 0; i < n; ++i) out << "1\n";
    out << "\n";
    out << "POINT_DATA " << n << "\n";
    out << "SCALARS diameter float\n";
    out << "LOOKUP_TABLE default \n";
    for (const object& o : objects) out << 2.0 * o.r << "\n";
    out << "\n";
}

// body centres with their diameters
void IO::exportParaviewBodies(const bodyList& bodies, const std::string& file) {
    std::ofstream out(file);
    out << "# vtk DataFile Version 3.0\n" << "Data.vtk\n" << "ASCII\n" << "DATASET POLYDATA\n" << "\n";
    out << "POINTS " << bodies.size() << " float\n";
    for (const body& b : bodies) b.x0.printFixedLine(out);
    out << "\n";
    out << "POINT_DATA " << bodies.size() << "\n";
    out << "SCALARS diameter float\n";
    out << "LOOKUP_TABLE default \n";
    for (const body& b : bodies) out << 2.0 * b.r << "\n";
    out << "\n";
}

// the spring as one line from the (first) body to the (first) impactor, with its compression
// L - L0 at both ends; the point count is 2 whatever the body and impactor counts, as written by
// the reference
void IO::exportParaviewSprings(const bodyList& bodies, const objectList& objects, const std::string& file) {
    std::ofstream out(file);
    out << "# vtk DataFile Version 3.0\n" << "Data.vtk\n" << "ASCII\n" << "DATASET POLYDATA\n" << "\n";
    out << "POINTS " << 2 << " float\n";
    for (const body& b : bodies) b.x0.printFixedLine(out);
    for (const object& o : objects) o.x0.printFixedLine(out);
    out << "\n";
    out << "LINES " << 1 << " " << 3 << "\n";
    out << "2 " << 0 << " " << 1 << "\n";
    out << "\n";
    out << "POINT_DATA " << 2 << "\n";
    out << "SCALARS radius float\n";
    out << "LOOKUP_TABLE default \n";
    out << "1.0\n" << "1.0\n" << "\n";
    out << "SCALARS compression float\n";
    out << "LOOKUP_TABLE default \n";
    for (const body& b : bodies) {
        out << b.L - b.L0 << "\n";
        out << b.L - b.L0 << "\n";
    }
    out << "\n";
}

void IO::exportBodyPosition(const bodyList& bodies) {
    std::ofstream f(bodyPositionFileName, std::ios::app);
    for (const body& b : bodies) {
        b.x0.print(f);
        f << "\t";
        b.x1.print(f);
        f << "\t" << b.L;
        f << std::endl;
    }
}

void IO::exportObjectForce(const objectList& objects, const LB& lb) {
    std::ofstream f(objectForceFileName, std::ios::app);
    for (const object& o : objects) {
        const double depth = (lb.fluidHeight + o.r) - o.x0.y;
        const tVect FExerted = o.FTotal - o.FGrav;
        tVect SExerted;
        if (depth > 0) {
            const double area = 4.0 * M_PI * o.r * o.r * (depth / (2.0 * o.r));
            SExerted = FExerted / area;
        }
        FExerted.print(f); f << "\t";
        SExerted.print(f); f << "\t";
        o.FTotal.print(f); f << "\t";
        o.FContact.print(f); f << "\t";
        o.FHydro.print(f); f << "\t";
        o.FLub.print(f); f << "\t";
        o.FRep.print(f);
        f << std::endl;
    }
}

void IO::exportWallForce(const wallList& walls) {
    std::ofstream f(wallForceFileName, std::ios::app);
    for (const wall& w : walls) {
        w.FTotal.print(f); f << "\t";
        w.FContact.print(f); f << "\t";
    }
    f << std::endl;
}

void IO::exportStressTensor(const DEM& dem) {
    std::ofstream f(stressTensorFileName, std::ios::app);
    dem.totalStress.print(f);
    f << std::endl;
}

void IO::exportNormalStressContributions(const DEM& dem) {
    std::ofstream f(normalStressContFileName, std::ios::app);
    dem.hydroStress.printNormal(f); f << "\t";
    dem.lubStress.printNormal(f); f << "\t";
    dem.contactStress.printNormal(f); f << "\t";
    dem.electroStress.printNormal(f);
    f << std::endl;
}

void IO::exportShearStressContributions(const DEM& dem) {
    std::ofstream f(shearStressContFileName, std::ios::app);
    dem.hydroStress.printShear(f); f << "\t";
    dem.lubStress.printShear(f); f << "\t";
    dem.contactStress.printShear(f); f << "\t";
    dem.electroStress.printShear(f);
    f << std::endl;
}

// radius, position, velocity, the negated symmetrised total stress and -disp.y of every particle
void IO::exportRecycleParticles(const particleList& particles, const std::string& file) {
    std::ofstream f(file);
    for (const particle& P : particles) {
        f << P.r << " ";
        P.x0.print(f);
        P.x1.print(f);
        const tMat& S = P.STotal;
        f << -S.m00 << "\t" << -S.m11 << "\t" << -S.m22 << "\t" << -(S.m01 + S.m10) * 0.5 << "\t"
          << -(S.m12 + S.m21) * 0.5 << "\t" << -(S.m02 + S.m20) * 0.5 << "\t" << -P.disp.y << "\n";
    }
}

void IO::exportMass(const LB& lb) {
    std::ofstream f(massFileName, std::ios::app);
    f.precision(17);
    f << currentTimeStep << " " << realTime << " " << lb.currentMass() << "\n";
}

// ------------------------------------------------------------------------------------------
// one coupled step (main.cpp:38-67)
// ------------------------------------------------------------------------------------------
void goCycle(IO& io, DEM& dem, LB& lb) {
    static const unsIntList boundary;  // the device holds the boundary codes
    io.realTime += lb.unitTime;
    ++io.currentTimeStep;
    ++lb.time;
    if (io.demSolve) dem.discreteElementStep(boundary);
    if (io.lbmSolve && dem.demTime >= dem.demInitialRepeat) {
        if (io.demSolve) lb.latticeBoltzmannCouplingStep(dem.newNeighborList, dem.particles, dem.objects);
        if (io.fluidVelocityDue()) lb.storeVelocity();
        lb.latticeBolzmannStep(dem.particles, dem.walls, dem.objects);
        if (lb.freeSurface) lb.latticeBoltzmannFreeSurfaceStep();
    } else if (io.lbmSolve && dem.demTime <= dem.demInitialRepeat) {
        if (io.demSolve) lb.latticeBoltzmannCouplingStep(dem.newNeighborList, dem.particles, dem.objects);
    }
    io.outputStep(lb, dem);
}

}  // namespace lbdem
