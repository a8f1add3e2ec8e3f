// design_setting.cpp -- getopt parsing with the reference's flags
// (3DgrainMPI/DesignSettingData.cpp:21-79): -f forcing folder prefix, -a full
// PF (not supported here), -s seed, -o output folder, -m MPI dims (ignored on
// one GPU), -d field saves, -l line configuration, -n 1|2 nucleation (1 = pure
// nucleation, liquid start), -b XYZ boundary digits.
#include "design_setting.h"

#include <getopt.h>

#include <cstdlib>
#include <iostream>

void DesignSettingData::getOptions(int argc, char** argv)
{
    if (argc < 2) {
        std::cerr << "usage: phase_field <input.py> [-s seed] [-b XYZ] [-o out] [-l 1] [-n 1|2] "
                     "[-d nsave] [-f forcing_prefix] [-k K] [-g 1] [-t type] [-G device]\n";
        std::exit(2);
    }
    inputFile = argv[1];
    static struct option long_options[] = {
        {"help", 0, 0, '?'},         {"macfile", 1, 0, 'f'},       {"useAPT", 1, 0, 'a'},
        {"seed", 1, 0, 's'},         {"output", 1, 0, 'o'},        {"mpiDim", 1, 0, 'm'},
        {"saveField", 1, 0, 'd'},    {"lineConfig", 1, 0, 'l'},    {"includeNuclean", 1, 0, 'n'},
        {"boundaryCondition", 1, 0, 'b'},                          {"slots", 1, 0, 'k'},
        {"gather27", 1, 0, 'g'},     {"thermalType", 1, 0, 't'},   {"device", 1, 0, 'G'},
        {"dump", 1, 0, 'D'},         {"parseOnly", 0, 0, 'P'},
        {0, 0, 0, 0}};
    optind = 2;
    int opt;
    while ((opt = getopt_long(argc, argv, "b:f:o:a:s:m:l:n:d:k:g:t:G:D:P?", long_options, NULL)) != -1) {
        switch (opt) {
        case 'f': thermalInputFolder = optarg; break;
        case 'a': useAPT = false; break;
        case 's': seedValue = std::atoi(optarg); break;
        case 'o': outputFolder = outputFolder + optarg; break;
        case 'm': mpiDim = std::atoi(optarg); break;
        case 'd': save3DField = std::atoi(optarg); break;
        case 'l': useLineConfig = true; break;
        case 'n':
            includeNucleation = true;
            if (std::atoi(optarg) == 1) pureNucleation = true;
            break;
        case 'b':
            givenBC = std::atoi(optarg);
            bcX = givenBC / 100;
            bcY = (givenBC - 100 * bcX) / 10;
            bcZ = givenBC % 10;
            break;
        case 'k': numSlots = std::atoi(optarg); break;
        case 'g': gather27 = std::atoi(optarg); break;
        case 't': thermalOverride = std::atoi(optarg); break;
        case 'G': device = std::atoi(optarg); break;
        case 'D': dumpDir = optarg; break;
        case 'P': parseOnly = true; break;
        default: break;
        }
    }
}
