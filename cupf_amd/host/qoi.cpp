// qoi.cpp -- see qoi.h.  Semantics follow 3DgrainMPI/QOI.cpp; ids of a
// junction are written in ascending order (the reference uses unordered_map
// iteration order, so only the set is meaningful), and index guards replace
// the reference's out-of-range writes for ids above NG.
#include "qoi.h"

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "../../include/pf_capi.h"

namespace {

// First plane above cur_tip whose interior holds a liquid point (alpha == 0),
// minus one (3DgrainMPI/QOI.cpp:47-67,162-185).
int scan_tip(int cur_tip, const int* alpha, int fnx, int fny, int fnz)
{
    for (;;) {
        ++cur_tip;
        if (cur_tip >= fnz - 1) break;
        const long long off = (long long)fnx * fny * cur_tip;
        bool liquid = false;
        for (int j = 1; j < fny - 1 && !liquid; ++j)
            for (int i = 1; i < fnx - 1; ++i)
                if (alpha[off + (long long)j * fnx + i] == 0) {
                    liquid = true;
                    break;
                }
        if (liquid) break;
    }
    return cur_tip - 1;
}

}  // namespace

QOILine::QOILine(const GlobalConstants& params)
{
    const size_t T = params.nts + 1, NG = params.num_theta;
    mQoIVectorFloatData.emplace("fractions", std::vector<float>(T * NG));
    mQoIVectorIntData.emplace("tip_y", std::vector<int>(T));
    mQoIVectorIntData.emplace("cross_sec", std::vector<int>());
    mQoIVectorIntData.emplace("extra_area", std::vector<int>(T * NG, 0));
    mQoIVectorIntData.emplace("total_area", std::vector<int>(T * NG, 0));
    mQoIVectorIntData.emplace("tip_final", std::vector<int>(T * NG));
    const int repeated_index = 5;
    mQoIVectorIntData.emplace("node_region",
                              std::vector<int>(T * repeated_index * params.num_nodes * mNumNodeFeatures, -1));
}

QOI3D::QOI3D(const GlobalConstants&)
{
    mQoIVectorIntData.emplace("volume", std::vector<int>());
    mQoIVectorIntData.emplace("grainToPF", std::vector<int>());
    mQoIVectorIntData.emplace("node_region", std::vector<int>());
    mQoIVectorIntData.emplace("manifold", std::vector<int>());
    mQoIVectorFloatData.emplace("geodesic_y_coors", std::vector<float>());
}

// 3DgrainMPI/QOI.cpp:38-97
void QOI::calculateLineQoIs(const GlobalConstants& params, int& cur_tip, const int* alpha, int kt,
                            const float* z, const int* loss_area, int move_count)
{
    const int fnx = params.fnx, fny = params.fny, fnz = params.fnz, NG = params.num_theta;
    cur_tip = scan_tip(cur_tip, alpha, fnx, fny, fnz);
    auto& tip_y = mQoIVectorIntData["tip_y"];
    if (kt >= 0 && kt < (int)tip_y.size()) tip_y[kt] = cur_tip + move_count;
    std::printf("frame %d, ntip %d, tip %f\n", kt, cur_tip + move_count, z[cur_tip]);
    auto& total = mQoIVectorIntData["total_area"];
    auto& extra = mQoIVectorIntData["extra_area"];
    for (int k = 1; k < fnz - 1; ++k)
        for (int j = 1; j < fny - 1; ++j)
            for (int i = 1; i < fnx - 1; ++i) {
                const int a = alpha[(long long)fnx * fny * k + (long long)fnx * j + i];
                if (a > 0 && a <= NG) {
                    total[(size_t)kt * NG + a - 1] += 1;
                    if (k > cur_tip) extra[(size_t)kt * NG + a - 1] += 1;
                }
            }
    for (int g = 0; g < NG; ++g) total[(size_t)kt * NG + g] += loss_area[g];
}

// 3DgrainMPI/QOI.cpp:198-231
void QOI::calculateQoIs(const GlobalConstants& params, const int* alpha, int)
{
    auto& volume = mQoIVectorIntData["volume"];
    auto& g2pf = mQoIVectorIntData["grainToPF"];
    volume.clear();
    g2pf.clear();
    std::unordered_map<int, int> slot;
    mNumActiveGrains = 0;
    const int fnx = params.fnx, fny = params.fny, fnz = params.fnz;
    for (int k = 1; k < fnz - 1; ++k)
        for (int j = 1; j < fny - 1; ++j)
            for (int i = 1; i < fnx - 1; ++i) {
                const int a = alpha[(long long)fnx * fny * k + (long long)fnx * j + i];
                if (a <= 0) continue;
                auto it = slot.find(a);
                if (it == slot.end()) {
                    it = slot.emplace(a, mNumActiveGrains++).first;
                    g2pf.push_back(a);
                    volume.push_back(0);
                }
                volume[it->second] += 1;
            }
}

bool QOI::calculateQoIsDevice(pf_ctx* ctx, const GlobalConstants& params)
{
    const int max_id = std::max(1, params.num_theta);
    std::vector<int> counts((size_t)max_id + 1), first((size_t)max_id + 1);
    long long above = 0;
    if (pf_qoi_volume(ctx, max_id, counts.data(), first.data(), &above) != PF_OK)
        throw std::runtime_error(std::string("pf_qoi_volume: ") + pf_last_error());
    if (above > 0) return false;
    std::vector<int> ids;
    for (int g = 1; g <= max_id; ++g)
        if (counts[g] > 0) ids.push_back(g);
    // first-seen order of the host scan (3DgrainMPI/QOI.cpp:198-231)
    std::sort(ids.begin(), ids.end(), [&](int a, int b) { return first[a] < first[b]; });
    auto& volume = mQoIVectorIntData["volume"];
    auto& g2pf = mQoIVectorIntData["grainToPF"];
    volume.clear();
    g2pf.clear();
    for (int g : ids) {
        g2pf.push_back(g);
        volume.push_back(counts[g]);
    }
    mNumActiveGrains = (int)ids.size();
    return true;
}

// 3DgrainMPI/QOI.cpp:162-185
void QOI::sampleHeights(int& cur_tip, const int* alpha, int fnx, int fny, int fnz)
{
    cur_tip = scan_tip(cur_tip, alpha, fnx, fny, fnz);
}

// 3DgrainMPI/QOI.cpp:38-97 with the tip scan and the per-grain counts on the
// device (pf_qoi_line); same results as calculateLineQoIs on the host alpha.
void QOI::calculateLineQoIsDevice(pf_ctx* ctx, const GlobalConstants& params, int& cur_tip, int kt,
                                  const float* z, const int* loss_area, int move_count)
{
    const int NG = params.num_theta;
    std::vector<int> t((size_t)NG), e((size_t)NG);
    int tip = cur_tip;
    if (pf_qoi_line(ctx, cur_tip, &tip, t.data(), e.data()) != PF_OK)
        throw std::runtime_error(std::string("pf_qoi_line: ") + pf_last_error());
    cur_tip = tip;
    auto& tip_y = mQoIVectorIntData["tip_y"];
    if (kt >= 0 && kt < (int)tip_y.size()) tip_y[kt] = cur_tip + move_count;
    std::printf("frame %d, ntip %d, tip %f\n", kt, cur_tip + move_count, z[cur_tip]);
    auto& total = mQoIVectorIntData["total_area"];
    auto& extra = mQoIVectorIntData["extra_area"];
    for (int g = 0; g < NG; ++g) {
        total[(size_t)kt * NG + g] += t[g] + loss_area[g];
        extra[(size_t)kt * NG + g] += e[g];
    }
}

// 3DgrainMPI/QOI.cpp:162-185 on the device alpha (pf_qoi_tip).
void QOI::sampleHeightsDevice(pf_ctx* ctx, int& cur_tip)
{
    int tip = cur_tip;
    if (pf_qoi_tip(ctx, cur_tip, &tip) != PF_OK)
        throw std::runtime_error(std::string("pf_qoi_tip: ") + pf_last_error());
    cur_tip = tip;
}

// 3DgrainMPI/QOI.cpp:103-160
void QOILine::searchJunctionsOnImage(const GlobalConstants& params, const int* alpha)
{
    const int fnx = params.fnx, fny = params.fny;
    auto& tip_y = mQoIVectorIntData["tip_y"];
    auto& cross = mQoIVectorIntData["cross_sec"];
    auto& nodes = mQoIVectorIntData["node_region"];
    const size_t per_frame = nodes.size() / (params.nts + 1);
    for (int kt = 0; kt < params.nts + 1; ++kt) {
        const long long off = (long long)fnx * fny * tip_y[kt];
        cross.insert(cross.end(), alpha + off, alpha + off + (long long)fnx * fny);
        size_t node_cnt = 0;
        for (int j = 1; j < fny - 1; ++j)
            for (int i = 1; i < fnx - 1; ++i) {
                std::vector<std::pair<int, int>> occ;  // (id, count)
                for (int dj = -1; dj <= 1; ++dj)
                    for (int di = -1; di <= 1; ++di) {
                        const int a = alpha[off + (long long)(j + dj) * fnx + i + di];
                        auto it = std::find_if(occ.begin(), occ.end(), [&](auto& e) { return e.first == a; });
                        if (it == occ.end()) occ.emplace_back(a, 1);
                        else ++it->second;
                    }
                int max_occur = 0;
                for (auto& e : occ) max_occur = std::max(max_occur, e.second);
                if (occ.size() >= 3 && max_occur <= 5) {
                    const size_t base = per_frame * kt + node_cnt * mNumNodeFeatures;
                    if (base + mNumNodeFeatures > per_frame * (kt + 1)) continue;
                    std::sort(occ.begin(), occ.end());
                    nodes[base] = i;
                    nodes[base + 1] = j;
                    nodes[base + 2] = max_occur;
                    for (size_t q = 0; q < occ.size() && 3 + q < (size_t)mNumNodeFeatures; ++q)
                        nodes[base + 3 + q] = occ[q].first;
                    ++node_cnt;
                }
            }
    }
}

// 3DgrainMPI/QOI.cpp:233-311
void QOI3D::searchJunctionsOnImage(const GlobalConstants& params, const int* alpha)
{
    auto& nodes = mQoIVectorIntData["node_region"];
    nodes.assign((size_t)40 * mNumActiveGrains * mNumNodeFeatures, -1);
    const int fnx = params.fnx, fny = params.fny, fnz = params.fnz;
    size_t node_cnt = 0;
    for (int k = 1; k < fnz - 1; ++k)
        for (int j = 1; j < fny - 1; ++j)
            for (int i = 1; i < fnx - 1; ++i) {
                const long long C = (long long)k * fnx * fny + (long long)j * fnx + i;
                if (alpha[C] == 0) continue;
                std::vector<std::pair<int, int>> occ;
                bool zero = false;
                for (int dk = -1; dk <= 1; ++dk)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            if (di * di + dj * dj + dk * dk > 2) continue;
                            const int a = alpha[(long long)(k + dk) * fnx * fny + (long long)(j + dj) * fnx + i + di];
                            if (a == 0) zero = true;
                            auto it = std::find_if(occ.begin(), occ.end(), [&](auto& e) { return e.first == a; });
                            if (it == occ.end()) occ.emplace_back(a, 1);
                            else ++it->second;
                        }
                if (zero) continue;
                int max_occur = 0;
                for (auto& e : occ) max_occur = std::max(max_occur, e.second);
                if (occ.size() == 4 && max_occur <= 10) {
                    const size_t base = node_cnt * mNumNodeFeatures;
                    if (base + mNumNodeFeatures > nodes.size()) continue;
                    std::sort(occ.begin(), occ.end());
                    nodes[base] = i;
                    nodes[base + 1] = j;
                    nodes[base + 2] = k;
                    nodes[base + 3] = max_occur;
                    for (size_t q = 0; q < 4; ++q) nodes[base + 4 + q] = occ[q].first;
                    ++node_cnt;
                }
            }
}

void QOI3D::searchJunctionsDevice(pf_ctx* ctx, const GlobalConstants&)
{
    auto& nodes = mQoIVectorIntData["node_region"];
    nodes.assign((size_t)40 * mNumActiveGrains * mNumNodeFeatures, -1);
    const int cap = (int)(nodes.size() / mNumNodeFeatures);
    std::vector<int> rows((size_t)8 * std::max(cap, 1));
    int n = 0;
    if (pf_qoi_junctions3d(ctx, rows.data(), cap, &n) != PF_OK)
        throw std::runtime_error(std::string("pf_qoi_junctions3d: ") + pf_last_error());
    // rows come in scan order; the host search keeps the first `cap` of them
    for (int q = 0; q < std::min(n, cap); ++q)
        for (int u = 0; u < 8; ++u) nodes[(size_t)q * mNumNodeFeatures + u] = rows[(size_t)8 * q + u];
}
