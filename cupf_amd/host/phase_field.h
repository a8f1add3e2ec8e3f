// phase_field.h -- drop-in for the reference's solver class surface
// (3DgrainMPI/cuPDE.h:4-21, PhaseField.h:11-59, APTPhaseField.h:11-30) on one
// GPU.  The public methods and members keep their names and roles; device
// work goes through the C ABI of include/pf_capi.h, and the MPI manager is a
// one-rank stand-in (no halo exchange: the physical boundary conditions are
// applied on the whole domain).
#pragma once

#include <string>
#include <vector>

#include "../../include/pf_capi.h"
#include "design_setting.h"
#include "params.h"
#include "qoi.h"

// One-rank stand-in for MPIsetting (3DgrainMPI/MPIsetting.h:12-63).
struct MPIsetting {
    int rank = 0, numProcessor = 1;
    int processorIDX = 0, processorIDY = 0, processorIDZ = 0;
    int numProcessorX = 1, numProcessorY = 1, numProcessorZ = 1;
    int nxLocal = 0, nyLocal = 0, nzLocal = 0, nzLocalAll = 0, haloWidth = 1;
};

// Macro thermal inputs and orientation tables (3DgrainMPI/ThermalInputData.h:8-25).
struct ThermalInputData {
    std::string folder;
    int Nx = 0, Ny = 0, Nz = 0, Nt = 0;
    std::vector<float> X_mac, Y_mac, Z_mac, t_mac, psi_mac, T_3D, theta_arr, cost, sint;
    std::vector<int> alpha_mac;
};

// Moving-window bookkeeping (3DgrainMPI/MovingDomain.h:3-18).
struct MovingDomain {
    int move_count = 0, cur_tip = 1, tip_front = 1, tip_thres = 0, samples = 0, lowsl = 1;
    std::vector<int> loss_area_host;
};

class PhaseField {
public:
    PhaseField() = default;
    virtual ~PhaseField();
    void parseInputParams(const std::string fileName);
    void cpuSetup(MPIsetting* mpiManager);
    void initField();
    virtual void cudaSetup();
    virtual void evolve();
    void OutputQoIs();
    void OutputField(int currentStep);
    // The GPU context holding the last collected alpha (none for the base class).
    virtual pf_ctx* deviceContext() const { return nullptr; }
    // Derived constants (params.txt, `name value` per line) and the initial
    // phi/alpha (raw float32/int32) of initField, for checking the set-up.
    void dumpSetup(const std::string& dir) const;
    void SetMPIManager(MPIsetting* m) { mMPIManager = m; }
    MPIsetting* GetMPIManager() const { return mMPIManager; }
    void SetDesignSetting(const DesignSettingData* d) { mDesignSetting = d; }
    const DesignSettingData* GetSetDesignSetting() const { return mDesignSetting; }

    int fnx = 0, fny = 0, fnz = 0, fnz_f = 0, NUM_PF = 8, length = 0, full_length = 0;
    std::vector<float> x, y, z, z_full, phi, psi;
    std::vector<int> alpha, alpha_i_full;
    ThermalInputData mac;
    GlobalConstants params;
    QOI* qois = nullptr;

protected:
    MPIsetting* mMPIManager = nullptr;
    const DesignSettingData* mDesignSetting = nullptr;
    std::string outputStem() const;
};

class APTPhaseField : public PhaseField {
public:
    APTPhaseField() = default;
    ~APTPhaseField() override;
    void cudaSetup() override;
    void evolve() override;
    pf_ctx* context() const { return ctx_; }
    pf_ctx* deviceContext() const override { return ctx_; }

protected:
    virtual void moveDomain(MovingDomain* movingDomainManager);
    virtual void getLineQoIs(MovingDomain* movingDomainManager);
    void setBC();

private:
    pf_ctx* ctx_ = nullptr;
    void check(int rc, const char* what) const;
};
