// phase_field.cpp -- host side of the drop-in (see phase_field.h).
//
// parseInputParams / cpuSetup / initField reproduce the reference's input
// contract and float arithmetic (3DgrainMPI/PhaseFieldCpu.cpp:17-459), so the
// same input files give the same grid, constants and initial field.
// APTPhaseField::evolve runs the reference's loop order and QoI / output
// cadence (3DgrainMPI/APTPhaseField.cu:710-946) over the C ABI.
#include "phase_field.h"

#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "h5io.h"

namespace {

constexpr double LS = -0.995;  // 3DgrainMPI/APTPhaseField.cu:25
constexpr int TIPP = 20;       // moving-domain check period (3DgrainMPI/APTPhaseField.cu:29)

// `key = value` on one line, word by word; later matches win and the value is
// read with stof even for ints (3DgrainMPI/helper.cpp:44-74).
template <class T>
void getParam(const std::string& line, const std::string& key, T& param)
{
    std::stringstream iss(line);
    std::string word;
    while (iss >> word) {
        if (word != key) continue;
        if (!(iss >> word) || word != "=") continue;
        if (!(iss >> word)) continue;
        try {
            param = (T)std::stof(word);
        } catch (...) {
        }
    }
}

// One value per line (3DgrainMPI/helper.cpp:77-101); returns the count read.
template <class T>
size_t read_input(const std::string& path, std::vector<T>& out, size_t n)
{
    out.assign(n, T(0));
    std::ifstream f(path);
    std::string line;
    size_t q = 0;
    while (q < n && std::getline(f, line)) {
        std::stringstream ss(line);
        double v = 0;
        if (ss >> v) out[q] = (T)v;
        ++q;
    }
    return q;
}

template <class T>
bool read_scalar(const std::string& path, T& v)
{
    std::vector<double> tmp;
    if (read_input(path, tmp, 1) == 0) return false;
    std::ifstream f(path);
    double d;
    if (!(f >> d)) return false;
    v = (T)d;
    return true;
}

std::string to_stringp(float a_value, int n)
{
    std::ostringstream out;
    out.precision(n);
    out << std::fixed << a_value;
    return out.str();
}

std::string basename_of(const std::string& p)
{
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? p : p.substr(s + 1);
}

}  // namespace

PhaseField::~PhaseField() { delete qois; }

void PhaseField::parseInputParams(const std::string fileName)
{
    const DesignSettingData* ds = GetSetDesignSetting();
    std::ifstream parseFile(fileName);
    if (!parseFile) throw std::runtime_error("cannot open input file " + fileName);
    std::string lineText;
    while (std::getline(parseFile, lineText)) {
        getParam(lineText, "Tmelt", params.Tmelt);
        getParam(lineText, "Dh", params.Dh);
        getParam(lineText, "L_cp", params.L_cp);
        getParam(lineText, "beta0", params.beta0);
        getParam(lineText, "GT", params.GT);
        getParam(lineText, "mu_k", params.mu_k);
        getParam(lineText, "kin_delta", params.kin_delta);
        getParam(lineText, "delta", params.delta);
        getParam(lineText, "underCoolingRate0", params.underCoolingRate0);
        getParam(lineText, "nuc_Nmax0", params.nuc_Nmax0);
        getParam(lineText, "undcool_mean", params.undcool_mean);
        getParam(lineText, "undcool_std", params.undcool_std);
        getParam(lineText, "nuc_rad", params.nuc_rad);
        getParam(lineText, "eps", params.eps);
        getParam(lineText, "dx", params.dx);
        getParam(lineText, "W0", params.W0);
        getParam(lineText, "asp_ratio_yx", params.asp_ratio_yx);
        getParam(lineText, "asp_ratio_zx", params.asp_ratio_zx);
        getParam(lineText, "moving_ratio", params.moving_ratio);
        getParam(lineText, "Lx", params.lxd);
        getParam(lineText, "eta", params.eta);
        getParam(lineText, "nts", params.nts);
        getParam(lineText, "ictype", params.ictype);
        getParam(lineText, "noi_period", params.noi_period);
        getParam(lineText, "haloWidth", params.haloWidth);
        getParam(lineText, "xmin", params.xmin);
        getParam(lineText, "ymin", params.ymin);
        getParam(lineText, "zmin", params.zmin);
        getParam(lineText, "nx", params.Nx);
        getParam(lineText, "ny", params.Ny);
        getParam(lineText, "nz", params.Nz);
        getParam(lineText, "nt", params.Nt);
        getParam(lineText, "cfl", params.cfl);
        getParam(lineText, "preMt", params.preMt);
    }
    if (params.haloWidth != 1) throw std::runtime_error("haloWidth must be 1 on one GPU");

    mac.Nx = params.Nx;
    mac.Ny = params.Ny;
    mac.Nz = params.Nz;
    mac.Nt = params.Nt;
    const std::string& F = mac.folder;
    read_input(F + "/x.txt", mac.X_mac, mac.Nx);
    read_input(F + "/y.txt", mac.Y_mac, mac.Ny);
    read_input(F + "/z.txt", mac.Z_mac, mac.Nz);
    read_input(F + "/t.txt", mac.t_mac, mac.Nt);
    read_input(F + "/psi.txt", mac.psi_mac, (size_t)mac.Nx * mac.Ny * mac.Nz);
    read_scalar(F + "/G.txt", params.G);
    read_scalar(F + "/Rmax.txt", params.R);
    read_scalar(F + "/V.txt", params.V);
    read_scalar(F + "/NG.txt", params.num_theta);
    read_scalar(F + "/NN.txt", params.num_nodes);
    read_scalar(F + "/UC.txt", params.underCoolingRate);
    read_scalar(F + "/Nmax.txt", params.nuc_Nmax);
    read_scalar(F + "/z0.txt", params.z0);
    read_scalar(F + "/r0.txt", params.r0);
    read_scalar(F + "/x0.txt", params.x0);
    read_scalar(F + "/top.txt", params.top);
    read_scalar(F + "/angle.txt", params.angle);
    read_scalar(F + "/min_angle.txt", params.min_angle);
    read_scalar(F + "/order.txt", params.order);
    params.nuc_Nmax0 = params.nuc_Nmax;
    if (params.num_theta < 1) throw std::runtime_error("NG.txt missing or < 1 in " + F);
    read_input(F + "/theta.txt", mac.theta_arr, 2 * (size_t)params.num_theta + 1);
    if (!h5io::read_f32(F + "/Temp.h5", "Temp", mac.T_3D)) {
        mac.T_3D.assign((size_t)mac.Nx * mac.Ny * mac.Nz * mac.Nt, 0.0f);
        std::cerr << "warning: " << F << "/Temp.h5 unreadable; T_3D set to 0\n";
    }

    // derived constants, in the reference's float arithmetic (:141-199)
    if (params.R < 0.3) {
        params.cfl *= 2;
        params.underCoolingRate0 /= 2;
    }
    if (params.R > 0.3 && params.R < 1) {
        params.cfl *= 1.5;
        params.underCoolingRate0 /= 1.5;
    }
    float dxd = params.dx * params.W0;
    params.c_infm = params.c_infty * params.m_slope;
    params.d0 = params.GT / params.L_cp;
    params.beta0 = 1.0 / (params.mu_k * params.L_cp);
    params.lamd = 0.8839 * params.W0 / params.d0;
    params.tau0 = params.beta0 * params.lamd * params.W0 / 0.8839;
    params.beta0_tilde = params.beta0 * params.W0 / params.tau0;
    params.dt = params.cfl * params.dx * params.beta0_tilde;
    params.hi = 1.0 / params.dx;
    params.sqrt2 = std::sqrt(2.0);
    params.dt_sqrt = std::sqrt(params.dt);
    params.nx = (int)std::round(params.lxd / params.dx / params.W0);
    params.ny = (int)std::round(params.asp_ratio_yx * params.nx);
    params.nz = ds->useLineConfig ? (int)std::round(params.moving_ratio * params.nx)
                                  : (int)std::round(params.asp_ratio_zx * params.nx);
    params.nz_full = (int)std::round(params.asp_ratio_zx * params.nx);
    params.lxd = params.nx * dxd;
    params.lyd = params.ny * dxd;
    params.lzd = params.nz * dxd;
    params.Mt = (int)(mac.t_mac[mac.Nt - 1] / params.tau0 / params.dt);
    params.Mt = (params.Mt / 2) * 2;
    int kts = params.Mt / params.nts;
    kts = (kts / 2) * 2;
    params.Mt = kts * params.nts;
    params.tmax = params.tau0 * params.dt * params.Mt;
    params.pts_cell = (int)(params.nuc_rad / dxd);
    params.bcX = ds->bcX;
    params.bcY = ds->bcY;
    params.bcZ = ds->bcZ;

    // thermal field chosen by the input file's name (:201-242), or forced with -t
    const std::string base = basename_of(ds->inputFile);
    if (base == "line.py") params.thermalType = PF_THERMAL_LINE;
    else if (base == "cylinder.py") params.thermalType = PF_THERMAL_CYLINDER;
    else if (base == "cone.py") params.thermalType = PF_THERMAL_CONE;
    else if (base == "paraboloid.py") {
        params.thermalType = PF_THERMAL_PARABOLOID;
        read_scalar(F + "/mp_len.txt", params.mp_len);
    } else params.thermalType = PF_THERMAL_NONE;
    if (ds->thermalOverride >= 0) params.thermalType = ds->thermalOverride;
    std::cout << "thermalType: " << params.thermalType << "  bc: " << params.bcX << params.bcY
              << params.bcZ << "  Mt: " << params.Mt << "  NG: " << params.num_theta << "\n";
}

void PhaseField::cpuSetup(MPIsetting* m)
{
    m->nxLocal = params.nx / m->numProcessorX;
    m->nyLocal = params.ny / m->numProcessorY;
    m->nzLocal = params.nz / m->numProcessorZ;
    m->nzLocalAll = params.nz_full;
    m->haloWidth = params.haloWidth;
    const float dxd = params.dx * params.W0;
    const float xmin_loc = params.xmin, ymin_loc = params.ymin, zmin_loc = params.zmin;
    m->nxLocal += 1;
    m->nyLocal += 1;
    m->nzLocal += 1;
    m->nzLocalAll += 1;
    params.fnx = m->nxLocal + 2 * params.haloWidth;
    params.fny = m->nyLocal + 2 * params.haloWidth;
    params.fnz = m->nzLocal + 2 * params.haloWidth;
    params.fnz_f = m->nzLocalAll + 2 * params.haloWidth;
    fnx = params.fnx;
    fny = params.fny;
    fnz = params.fnz;
    fnz_f = params.fnz_f;
    params.length = fnx * fny * fnz;
    params.full_length = fnx * fny * fnz_f;
    length = params.length;
    full_length = params.full_length;
    x.resize(fnx);
    y.resize(fny);
    z.resize(fnz);
    z_full.resize(fnz_f);
    for (int i = 0; i < fnx; i++) x[i] = (i - params.haloWidth) * dxd + xmin_loc;
    for (int i = 0; i < fny; i++) y[i] = (i - params.haloWidth) * dxd + ymin_loc;
    for (int i = 0; i < fnz; i++) z[i] = (i - params.haloWidth) * dxd + zmin_loc;
    for (int i = 0; i < fnz_f; i++) z_full[i] = (i - params.haloWidth) * dxd + zmin_loc;
    psi.assign(length, 0.0f);
    phi.assign(length, 0.0f);
    alpha.assign(length, 0);
    if (GetSetDesignSetting()->useLineConfig) alpha_i_full.assign(full_length, 0);
}

void PhaseField::initField()
{
    const DesignSettingData* ds = GetSetDesignSetting();
    const int NT = 2 * params.num_theta + 1;
    mac.sint.resize(NT);
    mac.cost.resize(NT);
    for (int i = 0; i < NT; i++) {
        mac.sint[i] = sinf(mac.theta_arr[i]);
        mac.cost[i] = cosf(mac.theta_arr[i]);
    }
    // Substrate ids from alpha.txt, extruded where phi > LS.  Applied whenever
    // the run is not pure nucleation and the file exists (the rule of
    // 3Dgrain/PhaseFieldCpu.cpp:389-436, which 3DgrainMPI restricts to -l).
    bool seed = false;
    if (!ds->pureNucleation) {
        struct stat st;
        seed = stat((mac.folder + "/alpha.txt").c_str(), &st) == 0;
        if (seed) read_input(mac.folder + "/alpha.txt", mac.alpha_mac, (size_t)(params.nx + 1) * (params.ny + 1));
    }
    const float Dx = mac.X_mac[mac.Nx - 1] - mac.X_mac[mac.Nx - 2];
    const float Dy = mac.Y_mac[mac.Ny - 1] - mac.Y_mac[mac.Ny - 2];
    const float Dz = mac.Z_mac[mac.Nz - 1] - mac.Z_mac[mac.Nz - 2];
    const int h = params.haloWidth;
    for (int id = 0; id < length; id++) {
        const int k = id / (fnx * fny);
        const int k_r = id - k * fnx * fny;
        const int j = k_r / fnx;
        const int i = k_r % fnx;
        if (i > h - 1 && i < fnx - h && j > h - 1 && j < fny - h && k > h - 1 && k < fnz - h) {
            int kx = (int)((x[i] - mac.X_mac[0]) / Dx);
            float delta_x = (x[i] - mac.X_mac[0]) / Dx - kx;
            int ky = (int)((y[j] - mac.Y_mac[0]) / Dy);
            float delta_y = (y[j] - mac.Y_mac[0]) / Dy - ky;
            int kz = (int)((z[k] - mac.Z_mac[0]) / Dz);
            float delta_z = (z[k] - mac.Z_mac[0]) / Dz - kz;
            if (kx == mac.Nx - 1) { kx = mac.Nx - 2; delta_x = 1.0f; }
            if (ky == mac.Ny - 1) { ky = mac.Ny - 2; delta_y = 1.0f; }
            if (kz == mac.Nz - 1) { kz = mac.Nz - 2; delta_z = 1.0f; }
            kx = std::min(std::max(kx, 0), mac.Nx - 2);
            ky = std::min(std::max(ky, 0), mac.Ny - 2);
            kz = std::min(std::max(kz, 0), mac.Nz - 2);
            const int off = kx + ky * mac.Nx + kz * mac.Nx * mac.Ny;
            const int offn = kx + ky * mac.Nx + (kz + 1) * mac.Nx * mac.Ny;
            const float* u = mac.psi_mac.data();
            psi[id] = ((1.0f - delta_x) * (1.0f - delta_y) * u[off] + (1.0f - delta_x) * delta_y * u[off + mac.Nx]
                       + delta_x * (1.0f - delta_y) * u[off + 1] + delta_x * delta_y * u[off + mac.Nx + 1]) * (1.0f - delta_z)
                    + ((1.0f - delta_x) * (1.0f - delta_y) * u[offn] + (1.0f - delta_x) * delta_y * u[offn + mac.Nx]
                       + delta_x * (1.0f - delta_y) * u[offn + 1] + delta_x * delta_y * u[offn + mac.Nx + 1]) * delta_z;
            psi[id] = psi[id] / params.W0;
            phi[id] = tanhf(psi[id] / params.sqrt2);
            if (phi[id] > LS && seed) {
                alpha[id] = mac.alpha_mac[(j - 1) * (fnx - 2 * h) + (i - 1)];
                if (alpha[id] < 1 || alpha[id] > params.num_theta)
                    std::cout << "found alpha out of bounds at " << (j - 1) * (fnx - 2 * h) + (i - 1)
                              << " alpha " << alpha[id] << "\n";
            } else {
                alpha[id] = 0;
            }
        } else {
            psi[id] = 0.0f;
            phi[id] = 0.0f;
            alpha[id] = 0;
        }
    }
    if (std::fabs(params.order - 100.0f) < 1e-3f) {
        std::cout << "overwrite psi field with static temperature profile\n";
        read_input(mac.folder + "/statTemp.txt", mac.psi_mac, (size_t)mac.Nx * mac.Ny * mac.Nz);
    }
}

void PhaseField::dumpSetup(const std::string& dir) const
{
    mkdir(dir.c_str(), 0755);
    std::ofstream f(dir + "/params.txt");
    f << std::setprecision(9);
#define P(name) f << #name << " " << params.name << "\n"
    P(nx); P(ny); P(nz); P(nz_full); P(fnx); P(fny); P(fnz); P(fnz_f); P(Mt); P(nts); P(preMt);
    P(dt); P(tau0); P(lamd); P(hi); P(cfl); P(d0); P(beta0); P(W0); P(dx); P(lxd); P(lyd); P(lzd);
    P(G); P(R); P(underCoolingRate); P(underCoolingRate0); P(z0); P(top); P(num_theta);
    P(thermalType); P(bcX); P(bcY); P(bcZ); P(pts_cell); P(kin_delta); P(L_cp); P(sqrt2);
    P(nuc_rad); P(undcool_mean); P(undcool_std); P(nuc_Nmax); P(nuc_Nmax0); P(seed_val);
#undef P
    // the macro grid as the kernels receive it (cudaSetup)
    f << "Nx " << mac.Nx << "\nNy " << mac.Ny << "\nNz " << mac.Nz << "\nNt " << mac.Nt << "\n";
    if (mac.Nx > 1 && mac.Ny > 0 && mac.Nz > 0 && mac.Nt > 0)
        f << "X0_mac " << mac.X_mac[0] << "\nY0_mac " << mac.Y_mac[0] << "\nZ0_mac " << mac.Z_mac[0]
          << "\nt0_mac " << mac.t_mac[0] << "\nDx_mac " << (mac.X_mac[1] - mac.X_mac[0]) << "\nDt_mac "
          << (mac.Nt > 1 ? mac.t_mac[1] - mac.t_mac[0] : 1.0f) << "\n";
    std::ofstream a(dir + "/init_phi.bin", std::ios::binary);
    a.write((const char*)phi.data(), sizeof(float) * phi.size());
    std::ofstream b(dir + "/init_alpha.bin", std::ios::binary);
    b.write((const char*)alpha.data(), sizeof(int) * alpha.size());
    std::ofstream c(dir + "/coords.bin", std::ios::binary);
    c.write((const char*)x.data(), sizeof(float) * x.size());
    c.write((const char*)y.data(), sizeof(float) * y.size());
    c.write((const char*)z.data(), sizeof(float) * z.size());
}

void PhaseField::cudaSetup() { throw std::runtime_error("the dense multi-phase-field solver is not part of this build; use APT"); }
void PhaseField::evolve() { throw std::runtime_error("the dense multi-phase-field solver is not part of this build; use APT"); }

std::string PhaseField::outputStem() const
{
    const DesignSettingData* ds = GetSetDesignSetting();
    const std::string grainType = ds->includeNucleation
        ? "Mixed_density" + std::to_string(params.nuc_Nmax) + "_grains" + std::to_string(qois->mNumActiveGrains)
        : "Epita_grains" + std::to_string(params.num_theta);
    return basename_of(ds->inputFile).substr(0, 4) + "seed" + std::to_string(params.seed_val) + "_lxd" +
           std::to_string(params.lxd) + "_G" + to_stringp(params.G, 3) + "_Rmax" + to_stringp(params.R, 3) +
           "_V" + to_stringp(params.V, 3) + "_angle" + to_stringp(params.angle, 3) + "_minangle" +
           to_stringp(params.min_angle, 3) + "_order" + std::to_string((int)params.order) + grainType;
}

void PhaseField::OutputQoIs()
{
    const DesignSettingData* ds = GetSetDesignSetting();
    if (ds->useLineConfig)
        qois->searchJunctionsOnImage(params, alpha_i_full.data());
    else if (deviceContext())  // the device copy of the last collected alpha: only the junctions cross PCIe
        static_cast<QOI3D*>(qois)->searchJunctionsDevice(deviceContext(), params);
    else
        qois->searchJunctionsOnImage(params, alpha.data());
    const std::string name = outputStem() + "_walltime" + std::to_string(params.walltime) + "_frames" +
                             std::to_string(params.nts) + "_Mt" + std::to_string(params.Mt) + "_rank" +
                             std::to_string(GetMPIManager()->rank) + ".h5";
    const std::string path = ds->outputFolder.empty() ? name : ds->outputFolder + "/" + name;
    std::cout << "save file name: " << path << "\n";
    h5io::Writer w(path);
    w.write_f32("x_coordinates", x.data(), fnx);
    w.write_f32("y_coordinates", y.data(), fny);
    w.write_f32("z_coordinates", z_full.data(), params.fnz_f);
    w.write_f32("angles", mac.theta_arr.data(), 2 * params.num_theta + 1);
    w.write_f32("r0", &params.r0, 1);
    w.write_f32("z0", &params.z0, 1);
    w.write_f32("x0", &params.x0, 1);
    w.write_f32("top", &params.top, 1);
    w.write_f32("angle", &params.angle, 1);
    w.write_f32("V", &params.V, 1);
    for (auto& q : qois->mQoIVectorFloatData) w.write_f32(q.first, q.second.data(), q.second.size());
    for (auto& q : qois->mQoIVectorIntData) w.write_i32(q.first, q.second.data(), q.second.size());
}

void PhaseField::OutputField(int currentStep)
{
    const DesignSettingData* ds = GetSetDesignSetting();
    const std::string name = outputStem() + "_frames" + std::to_string(params.nts) + "_Mt" +
                             std::to_string(params.Mt) + "_rank" + std::to_string(GetMPIManager()->rank) +
                             "_time" + std::to_string(currentStep) + ".h5";
    const std::string path = ds->outputFolder.empty() ? name : ds->outputFolder + "/" + name;
    std::cout << "save file name " << path << "\n";
    h5io::Writer w(path);
    w.write_f32("x_coordinates", x.data(), fnx);
    w.write_f32("y_coordinates", y.data(), fny);
    w.write_f32("z_coordinates", z_full.data(), params.fnz_f);
    w.write_f32("angles", mac.theta_arr.data(), 2 * params.num_theta + 1);
    if (ds->useLineConfig) w.write_i32("alpha", alpha_i_full.data(), full_length);
    else w.write_i32("alpha", alpha.data(), length);
}

// ------------------------------------------------------------- APT ---------
APTPhaseField::~APTPhaseField()
{
    if (ctx_) pf_destroy(ctx_);
}

void APTPhaseField::check(int rc, const char* what) const
{
    if (rc != PF_OK) throw std::runtime_error(std::string(what) + ": " + pf_last_error());
}

void APTPhaseField::cudaSetup()
{
    const DesignSettingData* ds = GetSetDesignSetting();
    params.NUM_PF = ds->numSlots;
    NUM_PF = params.NUM_PF;
    pf_params p{};
    p.fnx = fnx;
    p.fny = fny;
    p.fnz = fnz;
    p.K = NUM_PF;
    p.num_theta = params.num_theta;
    p.bcx = params.bcX;
    p.bcy = params.bcY;
    p.bcz = params.bcZ;
    p.thermal_type = params.thermalType;
    p.gather27 = ds->gather27;
    p.dt = params.dt;
    p.hi = params.hi;
    p.lamd = params.lamd;
    p.L_cp = params.L_cp;
    p.kin_delta = params.kin_delta;
    p.eps = params.eps;
    p.G = params.G;
    p.R = params.R;
    p.uc = params.underCoolingRate;
    p.uc0 = params.underCoolingRate0;
    p.z0 = params.z0;
    p.r0 = params.r0;
    p.lyd = params.lyd;
    p.lzd = params.lzd;
    p.V = params.V;
    p.mp_len = params.mp_len;
    p.x0 = params.x0;
    p.order = params.order;
    p.tan_angle = tanf(params.angle);
    p.cos_angle = cosf(params.angle);
    p.Nx = mac.Nx;
    p.Ny = mac.Ny;
    p.Nz = mac.Nz;
    p.Nt = mac.Nt;
    p.X0_mac = mac.X_mac[0];
    p.Y0_mac = mac.Y_mac[0];
    p.Z0_mac = mac.Z_mac[0];
    p.t0_mac = mac.t_mac[0];
    p.Dx_mac = mac.X_mac[1] - mac.X_mac[0];
    p.Dt_mac = mac.Nt > 1 ? mac.t_mac[1] - mac.t_mac[0] : 1.0f;
    p.pts_cell = (int)params.pts_cell;
    p.nuc_rad = params.nuc_rad;
    p.undcool_mean = params.undcool_mean;
    p.undcool_std = params.undcool_std;
    p.nuc_Nmax = params.nuc_Nmax;
    p.nuc_Nmax0 = params.nuc_Nmax0;
    p.dx = params.dx;
    p.sqrt2 = params.sqrt2;
    p.tau0 = params.tau0;
    p.seed = (uint64_t)(GetMPIManager()->rank + params.seed_val);
    check(pf_create(&p, ds->device, &ctx_), "pf_create");
    check(pf_set_coords(ctx_, x.data(), y.data(), z.data()), "pf_set_coords");
    std::vector<float> R9(9 * (size_t)(params.num_theta + 1));
    check(pf_orientations_from_angles(mac.theta_arr.data(), params.num_theta, R9.data()), "orientations");
    check(pf_set_orientations(ctx_, R9.data(), params.num_theta + 1), "pf_set_orientations");
    check(pf_set_thermal_macro(ctx_, mac.psi_mac.data(), mac.T_3D.empty() ? nullptr : mac.T_3D.data()),
          "pf_set_thermal_macro");
    check(pf_init_from_phi_alpha(ctx_, phi.data(), alpha.data()), "pf_init_from_phi_alpha");
}

void APTPhaseField::setBC() { check(pf_apply_bc(ctx_, 0), "pf_apply_bc"); }

void APTPhaseField::getLineQoIs(MovingDomain* md)
{
    const float loc = (md->move_count + md->lowsl - 1) * params.W0 * params.dx;
    const float threshold = params.z0 + params.top * (md->samples + 1) / params.nts;
    if (loc > threshold) {
        std::printf("threshold %f \n", threshold);
        // alpha stays on the device: the tip scan and per-grain counts run there
        check(pf_collect(ctx_, 0, PF_COLLECT_CLEAR_ALPHA, nullptr, nullptr), "pf_collect");
        check(pf_md_download(ctx_, nullptr, md->loss_area_host.data(), z.data()), "pf_md_download");
        md->cur_tip = 0;
        qois->calculateLineQoIsDevice(ctx_, params, md->cur_tip, md->samples + 1, z.data(),
                                      md->loss_area_host.data(), md->move_count);
        md->samples += 1;
    }
}

void APTPhaseField::moveDomain(MovingDomain* md)
{
    check(pf_md_tip(ctx_, &md->tip_front), "pf_md_tip");
    while (md->tip_front >= md->tip_thres) {
        check(pf_md_shift(ctx_, md->move_count), "pf_md_shift");
        md->move_count += 1;
        md->tip_front -= 1;
    }
    md->lowsl = 1;
    // the tip scan on the device alpha: fnz plane flags cross PCIe, not alpha
    check(pf_collect(ctx_, 0, 0, nullptr, nullptr), "pf_collect");
    qois->sampleHeightsDevice(ctx_, md->lowsl);
}

void APTPhaseField::evolve()
{
    const DesignSettingData* ds = GetSetDesignSetting();
    MovingDomain md;
    const int nucl = ds->includeNucleation ? PF_STEP_NUCLEATE : 0;
    if (ds->includeNucleation) check(pf_nucl_reset(ctx_, 1), "pf_nucl_reset");
    if (params.preMt > 0) {
        // liquid start: uniform cooling with nucleation, then re-initialise the
        // slots from the initial phi and the nucleated ids (:779-818)
        check(pf_nucl_reset(ctx_, 0), "pf_nucl_reset");
        for (int kt = 0; kt < params.preMt / 2; kt++)
            check(pf_step(ctx_, 2, 2 * kt + 1, PF_STEP_INITIAL_UC | PF_STEP_NUCLEATE), "pf_step");
        check(pf_collect(ctx_, 0, PF_COLLECT_CLEAR_ALPHA | PF_COLLECT_NO_BC, nullptr, nullptr), "pf_collect");
        check(pf_reinit_from_collect(ctx_), "pf_reinit_from_collect");
    }
    check(pf_collect(ctx_, 0, PF_COLLECT_CLEAR_ALPHA | PF_COLLECT_NO_BC, nullptr,
                     ds->useLineConfig ? nullptr : alpha.data()), "pf_collect");
    if (ds->useLineConfig) {
        md.tip_thres = (int)(0.8 * fnz);
        md.loss_area_host.assign(params.num_theta, 0);
        check(pf_md_enable(ctx_, fnz_f), "pf_md_enable");
        qois->calculateLineQoIsDevice(ctx_, params, md.cur_tip, 0, z.data(), md.loss_area_host.data(),
                                      md.move_count);
    } else if (!ds->includeNucleation) {
        qois->calculateQoIs(params, alpha.data(), 0);
    }
    const int qoikts = std::max(params.Mt / params.nts, 1);
    const int fieldkts = ds->save3DField > 0 ? std::max(params.Mt / ds->save3DField, 1) : 100000000;
    int kt_max = params.Mt / 2;
    if (ds->useLineConfig) kt_max += 1000;
    check(pf_sync(ctx_), "pf_sync");
    const auto t0 = std::chrono::steady_clock::now();
    int kt = 0;
    for (kt = 0; kt < kt_max; kt++) {
        check(pf_step(ctx_, 1, 2 * kt + 1, nucl), "pf_step");
        if (ds->useLineConfig) {
            getLineQoIs(&md);
            if (md.samples == params.nts) {
                std::printf("sample all %d heights\n", params.nts);
                break;
            }
            if ((2 * kt + 2) % TIPP == 0) moveDomain(&md);
        } else if ((2 * kt + 2) % qoikts == 0) {
            // per-grain volumes on the device; only the counts cross PCIe.  Ids
            // above num_theta (nucleated grains) take the host path.
            check(pf_collect(ctx_, 0, 0, nullptr, nullptr), "pf_collect");
            if (!qois->calculateQoIsDevice(ctx_, params)) {
                check(pf_collect(ctx_, 0, 0, nullptr, alpha.data()), "pf_collect");
                qois->calculateQoIs(params, alpha.data(), (2 * kt + 2) / qoikts);
            }
        }
        if ((2 * kt + 2) % fieldkts == 0) {
            check(pf_collect(ctx_, 0, PF_COLLECT_CLEAR_ALPHA, nullptr, alpha.data()), "pf_collect");
            if (ds->useLineConfig) {
                check(pf_md_download(ctx_, alpha_i_full.data(), nullptr, nullptr), "pf_md_download");
                const size_t off = (size_t)md.move_count * fnx * fny;
                const size_t n = std::min((size_t)length, alpha_i_full.size() - std::min(off, alpha_i_full.size()));
                std::copy(alpha.begin(), alpha.begin() + n, alpha_i_full.begin() + off);
            }
            OutputField(2 * kt + 2);
        }
        check(pf_step(ctx_, 1, 2 * kt + 2, 0), "pf_step");
    }
    check(pf_sync(ctx_), "pf_sync");
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    params.walltime = (int)sec;
    std::printf("time for %d iterations: %f s\n", 2 * kt, sec);
    long long ovf = 0;
    pf_overflow_count(ctx_, &ovf);
    std::printf("neighbourhoods over %d slots: %lld\n", NUM_PF, ovf);
    if (ds->includeNucleation) {
        long long nev = 0;
        pf_nucl_events(ctx_, &nev);
        std::printf("nucleation events: %lld\n", nev);
    }
    params.Mt = 2 * kt;
}
