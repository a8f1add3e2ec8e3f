// design_setting.h -- run options of the phase_field command line
// (3DgrainMPI/DesignSettingData.h:9-29), plus the device-side choices of this
// build (slots per point, gather neighbourhood, thermal-type override).
#pragma once
#include <string>

class DesignSettingData {
public:
    void getOptions(int argc, char** argv);

    bool useMPI = false;
    bool useLineConfig = false;
    bool includeNucleation = false;
    bool pureNucleation = false;
    bool useAPT = true;
    int save3DField = 1;
    int mpiDim = 1;
    int seedValue = 0;
    int givenBC = 0, bcX = 0, bcY = 0, bcZ = 0;
    std::string inputFile;
    std::string thermalInputFolder = "forcing/case";
    std::string outputFolder;
    // this build
    int numSlots = 8;          // -k: K active slots per point (5 or 8)
    int gather27 = 0;          // -g 1: 3Dgrain's 27-point id gather
    int thermalOverride = -1;  // -t: thermal type instead of the file-name rule
    int device = 0;            // -G: HIP device
    std::string dumpDir;       // -D: write derived constants and initial fields here
    bool parseOnly = false;    // -P: stop after initField (no device work)
};
