// Command-line driver mirroring the reference's main (src/mpulse.cpp:18-77): reads the setup
// file given as argv[1] (default "setup.dat"), builds the simulation and runs T-total steps on
// the GPU.  Options after the file name: --fp32 (FP32-bulk storage), --fp32res (float32 words
// with residual words), --reference (one kernel per reference loop), --sweep (single-pass
// sweep kernel), --pass (one x-march kernel per step: B of one step and E of the next).
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>

#include "mpulse_host.h"

int main(int argc, char **argv)
{
    std::string setupfilename = "setup.dat";
    if (argc > 1) setupfilename = argv[1];
    int precision = MPX_PREC_FP64, path = MPX_PATH_FUSED;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--fp32")) precision = MPX_PREC_FP32_BULK;
        else if (!std::strcmp(argv[a], "--fp32res")) precision = MPX_PREC_FP32_RES;
        else if (!std::strcmp(argv[a], "--pass")) path = MPX_PATH_PASS;
        else if (!std::strcmp(argv[a], "--reference")) path = MPX_PATH_REFERENCE;
        else if (!std::strcmp(argv[a], "--sweep")) path = MPX_PATH_SWEEP;
    }
    std::cerr << "****************************** This is MPulse (MI355X) *****************************\n";
    std::ifstream setupfile(setupfilename.c_str());
    std::cerr << "MAIN: Reading Inputfile: " << setupfilename << " ...\n";
    if (!setupfile) {
        std::cerr << "MAIN: Could not open: " << setupfilename << " ...\n";
        exit(-1);
    }
    try {
        mpulse_hip::Process process;
        process.RebuildAll(setupfile);
        for (const std::string &w : process.warnings) std::cerr << w << "\n";
        process.fieldSim->solver->precision = precision;
        process.fieldSim->solver->path = path;
        std::cerr << "                          *** MAIN:  INITIALIZING ***\n";
        process.init();
        std::cout << "                             *** MAIN: RUNNING ***\n";
        process.run();
    } catch (const mpulse_hip::SetupError &e) {
        std::cerr << e.msg << "\n";
        exit(-1);
    }
    std::cerr << "*************************** MPulse has finished ***************************\n";
    return 0;
}
