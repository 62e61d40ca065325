// ORACLE — test infrastructure only.
//
// Dump driver for the reference program compiled from its own sources where they lie
// (oracle/Makefile; nothing of the reference is copied). It runs the reference's
// initialisation sequence (main.cpp:262-286) and goCycle (main.cpp:38-67) through the
// reference's own functions, and after selected steps writes the solver state that the
// reference keeps in public members (LB.h:28-150, DEM.h:22-118) to a binary file whose
// layout matches the records of include/lbdem.h. Output is used to write golden fixtures
// and to pin the FP64 restatement in oracle/refcpu.cpp.
//
// usage: ref_driver <config.cfg> <workdir> <nsteps> <dump-prefix> [step ...]
//        (run from the directory holding input/particles.dat, input/objects.dat, input/body.dat)
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <set>
#include <string>
#include <vector>

#include <omp.h>

#include "DEM.h"
#include "IO.h"
#include "LB.h"
#include "getpot.h"
#include "../include/lbdem.h"

// defined in the reference's main.cpp (compiled with -Dmain=lbdem_reference_main)
void goCycle(IO& io, DEM& dem, LB& lb);
void parseCommandLine(IO& io, GetPot& commandLine);
void parseConfigFile(IO& io, DEM& dem, LB& lb, GetPot& configFile, GetPot& commandLine);

static void put3(double* d, const tVect& v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

static void dump(const std::string& path, LB& lb, DEM& dem, IO& io) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) { std::perror(path.c_str()); std::exit(3); }
    const int64_t N = lb.totNodes;
    int32_t hdr[8] = {int32_t(lb.lbSize[0]), int32_t(lb.lbSize[1]), int32_t(lb.lbSize[2]), int32_t(dem.particles.size()),
                      int32_t(dem.objects.size()), int32_t(dem.walls.size()), int32_t(lb.time), int32_t(dem.demTimeStep)};
    std::fwrite("LBDMDUMP", 1, 8, fp);
    std::fwrite(hdr, sizeof(hdr), 1, fp);
    std::vector<uint8_t> t(N), p(N), ex(N);
    std::vector<int32_t> s(N);
    std::vector<double> f(size_t(N) * 19, 0.0), n(N, 0.0), m(N, 0.0);
    for (int64_t i = 0; i < N; ++i) {
        t[i] = uint8_t(lb.types[i].getType());
        p[i] = lb.types[i].isInsideParticle() ? 1 : 0;
        s[i] = int32_t(lb.types[i].getSolidIndex());
        ex[i] = lb.nodes[i] != 0;
        if (lb.nodes[i]) {
            for (int q = 0; q < 19; ++q) f[size_t(q) * N + i] = lb.nodes[i]->f[q];
            n[i] = lb.nodes[i]->n;
            m[i] = lb.nodes[i]->mass;
        }
    }
    std::fwrite(t.data(), 1, N, fp);
    std::fwrite(p.data(), 1, N, fp);
    std::fwrite(s.data(), 4, N, fp);
    std::fwrite(ex.data(), 1, N, fp);
    std::fwrite(f.data(), 8, f.size(), fp);
    std::fwrite(n.data(), 8, N, fp);
    std::fwrite(m.data(), 8, N, fp);
    for (size_t i = 0; i < dem.particles.size(); ++i) {
        const particle& P = dem.particles[i];
        double r[LBDEM_P_REC] = {0};
        put3(r + LBDEM_P_X0, P.x0); put3(r + LBDEM_P_X1, P.x1); put3(r + LBDEM_P_W1, P.w1);
        put3(r + LBDEM_P_FHYDRO, P.FHydro); put3(r + LBDEM_P_MHYDRO, P.MHydro);
        put3(r + LBDEM_P_FCONTACT, P.FContact); put3(r + LBDEM_P_MCONTACT, P.MContact);
        put3(r + LBDEM_P_FLUB, P.FLub); put3(r + LBDEM_P_MLUB, P.MLub); put3(r + LBDEM_P_FREP, P.FRep);
        put3(r + LBDEM_P_FTOTAL, P.FTotal); put3(r + LBDEM_P_MTOTAL, P.MTotal); put3(r + LBDEM_P_FGRAV, P.FGrav);
        r[LBDEM_P_FLUIDVOL] = P.fluidVolume;
        put3(r + LBDEM_P_XIW, P.xiw); put3(r + LBDEM_P_XIO, P.xio);
        r[LBDEM_P_R] = P.r; r[LBDEM_P_M] = P.m;
        std::fwrite(r, 8, LBDEM_P_REC, fp);
    }
    for (size_t i = 0; i < dem.objects.size(); ++i) {
        const object& O = dem.objects[i];
        double r[LBDEM_O_REC] = {0};
        put3(r + LBDEM_O_X0, O.x0); put3(r + LBDEM_O_X1, O.x1); put3(r + LBDEM_O_W1, O.w1);
        put3(r + LBDEM_O_FHYDRO, O.FHydro); put3(r + LBDEM_O_MHYDRO, O.MHydro);
        put3(r + LBDEM_O_FCONTACT, O.FContact); put3(r + LBDEM_O_MCONTACT, O.MContact);
        put3(r + LBDEM_O_FLUB, O.FLub); put3(r + LBDEM_O_MLUB, O.MLub);
        put3(r + LBDEM_O_FTOTAL, O.FTotal); put3(r + LBDEM_O_MTOTAL, O.MTotal); put3(r + LBDEM_O_FGRAV, O.FGrav);
        r[LBDEM_O_R] = O.r; r[LBDEM_O_M] = O.m;
        std::fwrite(r, 8, LBDEM_O_REC, fp);
    }
    for (size_t i = 0; i < dem.walls.size(); ++i) {
        const wall& W = dem.walls[i];
        double r[LBDEM_W_REC] = {0};
        put3(r + LBDEM_W_FHYDRO, W.FHydro); put3(r + LBDEM_W_FCONTACT, W.FContact);
        put3(r + LBDEM_W_FLUB, W.FLub); put3(r + LBDEM_W_FTOTAL, W.FTotal);
        std::fwrite(r, 8, LBDEM_W_REC, fp);
    }
    const int cap = 64;
    std::vector<int32_t> cnt(dem.particles.size()), lst(dem.particles.size() * cap, -1);
    for (size_t i = 0; i < dem.particles.size(); ++i) {
        const particle& P = dem.particles[i];
        std::vector<int> v(P.list + 1, P.list + 1 + P.list[0]);
        std::sort(v.begin(), v.end());
        cnt[i] = int32_t(v.size());
        for (size_t k = 0; k < v.size() && k < size_t(cap); ++k) lst[i * cap + k] = v[k];
    }
    std::fwrite(cnt.data(), 4, cnt.size(), fp);
    std::fwrite(lst.data(), 4, lst.size(), fp);
    // total fluid mass with the definition of LB.cpp:388-395, over the current active list
    double tm = 0.0;
    for (size_t k = 0; k < lb.activeNodes.size(); ++k) {
        const unsigned idx = lb.activeNodes[k];
        if (!lb.types[idx].isInsideParticle() && lb.nodes[idx]) tm += lb.nodes[idx]->mass;
    }
    double sc[6] = {lb.totalMass, tm, dem.demTime, io.realTime, double(lb.time), double(dem.demTimeStep)};
    std::fwrite(sc, 8, 6, fp);
    // bodies (elmt.h body): count, then LBDEM_B_REC values each
    const int32_t nb = int32_t(dem.bodies.size());
    std::fwrite(&nb, 4, 1, fp);
    for (size_t i = 0; i < dem.bodies.size(); ++i) {
        const body& B = dem.bodies[i];
        double r[LBDEM_B_REC] = {};
        const tVect* v[5] = {&B.x0, &B.x1, &B.FSpring, &B.FTotal, &B.FGrav};
        const int at[5] = {LBDEM_B_X0, LBDEM_B_X1, LBDEM_B_FSPRING, LBDEM_B_FTOTAL, LBDEM_B_FGRAV};
        for (int q = 0; q < 5; ++q) { r[at[q]] = v[q]->x; r[at[q] + 1] = v[q]->y; r[at[q] + 2] = v[q]->z; }
        r[LBDEM_B_L] = B.L; r[LBDEM_B_L0] = B.L0; r[LBDEM_B_K] = B.k; r[LBDEM_B_R] = B.r; r[LBDEM_B_M] = B.m;
        std::fwrite(r, 8, LBDEM_B_REC, fp);
    }
    // stress diagnostics in the layout of lbdem_get_stress: 45 global values, then STotal (9) and
    // disp (3) of every particle
    const tMat* g[5] = {&dem.totalStress, &dem.hydroStress, &dem.lubStress, &dem.contactStress, &dem.electroStress};
    for (const tMat* t : g) {
        const double m[9] = {t->m00, t->m01, t->m02, t->m10, t->m11, t->m12, t->m20, t->m21, t->m22};
        std::fwrite(m, 8, 9, fp);
    }
    for (size_t i = 0; i < dem.particles.size(); ++i) {
        const tMat& t = dem.particles[i].STotal;
        const double m[9] = {t.m00, t.m01, t.m02, t.m10, t.m11, t.m12, t.m20, t.m21, t.m22};
        std::fwrite(m, 8, 9, fp);
    }
    for (size_t i = 0; i < dem.particles.size(); ++i) {
        const tVect& d = dem.particles[i].disp;
        const double v[3] = {d.x, d.y, d.z};
        std::fwrite(v, 8, 3, fp);
    }
    // the rest of the particle VTK file's values, in the layout of lbdem_get_particle_stress
    for (size_t i = 0; i < dem.particles.size(); ++i) {
        const particle& P = dem.particles[i];
        double r[LBDEM_S_REC] = {};
        const tMat* t[4] = {&P.STotal, &P.SElas, &P.SVisc, &P.SVisc_h};
        for (int k = 0; k < 4; ++k) {
            const double m[9] = {t[k]->m00, t[k]->m01, t[k]->m02, t[k]->m10, t[k]->m11, t[k]->m12, t[k]->m20, t[k]->m21, t[k]->m22};
            for (int e = 0; e < 9; ++e) r[9 * k + e] = m[e];
        }
        r[LBDEM_S_OVERLAP] = P.overlap.x; r[LBDEM_S_OVERLAP + 1] = P.overlap.y; r[LBDEM_S_OVERLAP + 2] = P.overlap.z;
        r[LBDEM_S_DISP] = P.disp.x; r[LBDEM_S_DISP + 1] = P.disp.y; r[LBDEM_S_DISP + 2] = P.disp.z;
        r[LBDEM_S_COORD] = P.coord; r[LBDEM_S_LUB_COORD] = P.lub_coord;
        std::fwrite(r, 8, LBDEM_S_REC, fp);
    }
    std::fclose(fp);
}

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <config.cfg> <workdir> <nsteps> <dump-prefix> [step ...]\n", argv[0]);
        return 2;
    }
    const std::string cfg = argv[1], work = argv[2], prefix = argv[4];
    const int nsteps = std::atoi(argv[3]);
    std::set<int> steps;
    for (int i = 5; i < argc; ++i) steps.insert(std::atoi(argv[i]));

    IO io;
    DEM dem;
    LB lb;
    time_t tt = time(0);
    io.now = localtime(&tt);
    std::vector<std::string> args = {"ref_driver", "-c", cfg, "-d", work};
    std::vector<char*> av;
    for (auto& a : args) av.push_back(const_cast<char*>(a.c_str()));
    GetPot commandLine(int(av.size()), av.data());
    parseCommandLine(io, commandLine);
    GetPot configFile(io.configFileName);
    parseConfigFile(io, dem, lb, configFile, commandLine);
    io.currentTimeStep = 0;
    io.initialize();
    lb.latticeDefinition();
    const tVect externalForce = lb.lbF * lb.unit.Accel;
    dem.discreteElementInit(lb.boundary, lb.lbPhysicalSize, lb.lbBoundaryLocation, lb.unit.Time, externalForce,
                            lb.fluidHeight, lb.freeSurface);
    if (io.lbmSolve) lb.latticeBolzmannInit(dem.walls, dem.particles, dem.objects);
    lb.time = 0;
    dem.demTime = 0.0;
    io.realTime = 0.0;
    dem.demTimeStep = 0;
    io.outputStep(lb, dem);
    if (steps.count(0)) dump(prefix + "_0.bin", lb, dem, io);
    const double t0 = omp_get_wtime();
    double tWarm = t0;
    const int warm = std::getenv("LBDEM_WARMUP") ? std::atoi(std::getenv("LBDEM_WARMUP")) : 0;
    for (int s = 1; s <= nsteps; ++s) {
        goCycle(io, dem, lb);
        if (s == warm) tWarm = omp_get_wtime();
        if (steps.count(s)) dump(prefix + "_" + std::to_string(s) + ".bin", lb, dem, io);
    }
    const double t1 = omp_get_wtime();
    std::printf("LBDEM_TIMING steps=%d warmup=%d seconds=%.6f active=%zu threads=%d\n", nsteps - warm, warm, t1 - tWarm,
                lb.activeNodes.size(), omp_get_max_threads());
    return 0;
}
