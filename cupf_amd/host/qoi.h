// qoi.h -- host quantities of interest computed from the grain-id field alpha
// (3DgrainMPI/QOI.h:12-47): per-grain volume, quadruple-junction search,
// line-mode tip height / areas / cross-sections, moving-window tip samples.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "params.h"

struct pf_ctx;

class QOI {
public:
    virtual ~QOI() = default;
    virtual void searchJunctionsOnImage(const GlobalConstants& params, const int* alpha) = 0;
    void calculateLineQoIs(const GlobalConstants& params, int& cur_tip, const int* alpha, int kt,
                           const float* z, const int* loss_area, int move_count);
    void calculateQoIs(const GlobalConstants& params, const int* alpha, int kt);
    // calculateQoIs on the context's device alpha (the last pf_collect);
    // false (nothing changed) if some id exceeds num_theta (nucleated grains):
    // the caller then downloads alpha and uses calculateQoIs.
    bool calculateQoIsDevice(pf_ctx* ctx, const GlobalConstants& params);
    void sampleHeights(int& cur_tip, const int* alpha, int fnx, int fny, int fnz);
    // calculateLineQoIs and sampleHeights on the context's device alpha (the
    // last pf_collect): only plane flags and per-grain counts cross PCIe.
    void calculateLineQoIsDevice(pf_ctx* ctx, const GlobalConstants& params, int& cur_tip, int kt,
                                 const float* z, const int* loss_area, int move_count);
    void sampleHeightsDevice(pf_ctx* ctx, int& cur_tip);
    std::map<std::string, std::vector<int>> mQoIVectorIntData;
    std::map<std::string, std::vector<float>> mQoIVectorFloatData;
    int mNumActiveGrains = 0;
};

class QOILine : public QOI {
public:
    explicit QOILine(const GlobalConstants& params);
    void searchJunctionsOnImage(const GlobalConstants& params, const int* alpha) override;

private:
    int mNumNodeFeatures = 8;
};

class QOI3D : public QOI {
public:
    explicit QOI3D(const GlobalConstants& params);
    void searchJunctionsOnImage(const GlobalConstants& params, const int* alpha) override;
    // the same search on the context's device alpha (the last pf_collect)
    void searchJunctionsDevice(pf_ctx* ctx, const GlobalConstants& params);

private:
    int mNumNodeFeatures = 9;
};
