// lbdem_run — the reference program's main loop (main.cpp:207-310) over the C ABI.
//
//   lbdem_run -c config.cfg -d workdir [-n name] [-<key> <value> ...]
//             [--library path/to/liblbdem.so] [--precision fp64|fp32|fp16] [--device N]
//
// Reads ./input/particles.dat and ./input/objects.dat relative to the working directory, as the
// reference does, runs goCycle until realTime > maxTime and writes export.dat, objPosition.dat,
// objForce.dat, wallForce.dat and mass.dat into the work directory on the screenExpTime interval.
// The default library is lib/liblbdem.so next to this program's bin/ directory.
#include <sys/stat.h>
#include <unistd.h>

#include <climits>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "facade.hpp"

using namespace lbdem;

static std::string default_library() {
    char buf[PATH_MAX];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (n <= 0) return "liblbdem.so";
    buf[n] = 0;
    std::string exe(buf);
    const std::string dir = exe.substr(0, exe.find_last_of('/'));
    return dir + "/../lib/liblbdem.so";
}

static void usage() {
    std::cout << "usage: lbdem_run -c config.cfg -d workdir [-n name] [-key value ...]\n"
                 "                 [--library lib.so] [--precision fp64|fp32|fp16] [--device N]\n";
}

int main(int argc, char** argv) {
    CommandLine cl(argc, argv);
    if (cl.search("-h") || cl.search("--help")) {
        usage();
        return 0;
    }
    std::string library = default_library();
    if (cl.search("--library")) library = cl.next(library);
    int precision = LBDEM_PREC_FP64;
    if (cl.search("--precision")) {
        const std::string p = cl.next("fp64");
        precision = p == "fp32" ? LBDEM_PREC_FP32 : (p == "fp16" ? LBDEM_PREC_FP16 : LBDEM_PREC_FP64);
    }
    int device = 0;
    if (cl.search("--device")) device = cl.next(0);

    try {
        Context ctx(library, precision, device);
        IO io(ctx);
        DEM dem(ctx);
        LB lb(ctx);

        // command line (main.cpp:70-116)
        io.workDirectory = "data";
        if (cl.search("-d")) io.workDirectory = cl.next(io.workDirectory);
        if (cl.search("-n")) {
            const std::string name = cl.next("name");
            if (name != "time") io.workDirectory = io.workDirectory + "/" + name;
        }
        ::mkdir(io.workDirectory.c_str(), 0777);
        io.configFileName = "config.cfg";
        if (cl.search("-c")) {
            io.configFileName = cl.next(io.configFileName);
            std::cout << "Using " << io.configFileName << " for LBM and DEM parameters\n";
            std::ifstream src(io.configFileName, std::ios::binary);
            const std::string base = io.configFileName.substr(io.configFileName.find_last_of('/') + 1);
            std::ofstream dst(io.workDirectory + "/" + base, std::ios::binary);
            dst << src.rdbuf();
        }
        ConfigFile cfg(io.configFileName);
        if (!cfg.ok()) std::cout << "ERROR: Can't open config file \"" << io.configFileName << "\" for reading!\n";
        parseConfigFile(io, dem, lb, cfg, cl);

        // initialisation (main.cpp:262-286)
        io.initialize();
        lb.latticeDefinition();
        lb.LBShow();
        bool freeSurface = lb.freeSurface != 0;
        dem.discreteElementInit(unsIntList(), doubleList(), vecList(), lb.unitTime, tVect(), lb.fluidHeight,
                                freeSurface);
        if (io.lbmSolve) lb.latticeBolzmannInit(dem.walls, dem.particles, dem.objects);
        lb.time = 0;
        dem.demTime = 0.0;
        io.realTime = 0.0;
        dem.demTimeStep = 0;
        io.outputStep(lb, dem);

        // time loop (main.cpp:288-308)
        while (!(io.realTime != 0.0 && io.realTime > io.maxTime)) goCycle(io, dem, lb);
    } catch (const std::exception& e) {
        std::cout << "ERROR: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
