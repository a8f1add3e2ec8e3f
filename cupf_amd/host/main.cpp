// main.cpp -- phase_field command line, the drop-in for the reference's
// driver (3DgrainMPI/main.cpp:15-77) on one GPU:
//   phase_field <input.py> -s SEED -b XYZ -o OUT [-l 1] [-n 1|2] [-d N] [-f PREFIX] [-k K] [-g 1]
#include <iostream>
#include <string>

#include "phase_field.h"

int main(int argc, char** argv)
{
    DesignSettingData designSetting;
    designSetting.getOptions(argc, argv);
    if (!designSetting.useAPT) {
        std::cerr << "phase_field: only the active-parameter-tracking solver is built (drop -a)\n";
        return 2;
    }
    MPIsetting mpiManager;
    try {
        APTPhaseField solver;
        solver.SetDesignSetting(&designSetting);
        solver.SetMPIManager(&mpiManager);
        solver.mac.folder = designSetting.thermalInputFolder + std::to_string(designSetting.seedValue);
        solver.params.seed_val = designSetting.seedValue;
        solver.parseInputParams(designSetting.inputFile);
        solver.cpuSetup(&mpiManager);
        if (designSetting.useLineConfig) solver.qois = new QOILine(solver.params);
        else solver.qois = new QOI3D(solver.params);
        std::cout << "field initialization on cpu\n";
        solver.initField();
        if (!designSetting.dumpDir.empty()) solver.dumpSetup(designSetting.dumpDir);
        if (designSetting.parseOnly) return 0;
        std::cout << "device setup: " << pf_build_info() << "\n";
        solver.cudaSetup();
        solver.evolve();
        solver.OutputQoIs();
    } catch (const std::exception& e) {
        std::cerr << "phase_field: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
