"""The host check that runs before every step launch (cupf_amd/csrc/capacity.h,
used by check_capacity in pf_capi.hip): its worst-case list sizes against a
brute-force enumeration of what the plan kernel can list, compiled with g++
on the CPU.  A list sized from the wrong factors (the class of bug that once
sized the work list to one entry) makes check_capacity refuse the launch."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <set>
#include <utility>
#include "capacity.h"
using namespace cupf;
// (tile, z chunk) pairs holding at least one interior point outside the
// narrow +x edge strip, found point by point
static long long brute(int fnx, int fny, int fnz, int kchunk, int pmin) {
    const int n = fnx - 2, full = n / 64, ex = n - 64 * full;
    const int tx_lim = ex >= pmin ? full + 1 : full;   // tiles of the LDS path in x
    std::set<std::pair<long long, int>> seen;
    for (int k = 1; k <= fnz - 2; ++k)
        for (int j = 1; j <= fny - 2; ++j)
            for (int i = 1; i <= fnx - 2; ++i) {
                const int tx = (i - 1) / 64, ty = (j - 1) / 4;
                if (tx >= tx_lim) continue;            // edge strip: computed items
                seen.insert({(long long)ty * 100000 + tx, (k - 1) / kchunk});
            }
    return (long long)seen.size();
}
int main() {
    const int grids[][3] = {{5, 5, 5}, {19, 12, 30}, {67, 67, 67}, {66, 9, 40}, {82, 81, 20},
                            {130, 11, 17}, {259, 13, 70}, {34, 34, 34}, {98, 98, 130}};
    int bad = 0, n = 0;
    for (auto& g : grids)
        for (int kc = 1; kc <= 61; kc += (kc < 8 ? 1 : 9)) {
            const long long w = worst_list_items(g[0], g[1], g[2], kc, 16);
            const long long b = brute(g[0], g[1], g[2], kc, 16);
            ++n;
            if (w != b) { std::printf("grid %d %d %d kchunk %d: %lld vs %lld\n", g[0], g[1], g[2], kc, w, b); ++bad; }
            // the batch list, sized at depth 2, holds every depth >= 2
            if (kc >= 2 && w > worst_list_items(g[0], g[1], g[2], 2, 16)) ++bad;
            if (worst_fix_entries(g[0], g[1], g[2]) != (long long)(g[0] - 2) * (g[1] - 2) * (g[2] - 2)) ++bad;
        }
    std::printf("%d cases, %d mismatches\n", n, bad);
    return bad != 0;
}
"""


def test_worst_case_list_sizes_match_brute_force(tmp_path):
    src = tmp_path / "cap.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "cap"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "cupf_amd", "csrc"), str(src),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "0 mismatches" in r.stdout


def test_every_launch_path_checks_capacity():
    """Both launch paths call the check before enqueueing work."""
    src = open(os.path.join(ROOT, "cupf_amd", "csrc", "pf_capi.hip")).read()
    step = src[src.index("static int step_impl("):]
    step = step[:step.index("launch_step(")]
    assert "check_capacity(c)" in step
    batch = src[src.index("int pf_batch_step("):]
    batch = batch[:batch.index("batch_run(")]
    assert "check_capacity(c)" in batch and "worst_list_items" in batch
