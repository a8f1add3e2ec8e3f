// capacity.h -- worst-case entries of a step's work lists, recomputed from
// the grid and the chunk depth alone (not from the values pf_create derived
// and stored), for the host check before every launch.  Plain C++: the CPU
// tests compile it on its own.
#pragma once

namespace cupf {

// The LDS path's column tiles are 64 x 4 interior points; a leftover strip at
// the +x end at least partial_min wide is a partial tile, a narrower one goes
// to edge items (which are computed, never listed).
inline int cap_tiles_x(int fnx, int partial_min)
{
    const int n = fnx - 2, t = n / 64;
    return (n - 64 * t >= partial_min) ? t + 1 : t;
}

inline int cap_tiles_y(int fny) { return (fny - 2 + 3) / 4; }

inline int cap_nzchunk(int fnz, int kchunk) { return (fnz - 2 + kchunk - 1) / kchunk; }

// Entries the plan kernel may list in one step: each (column tile, z chunk)
// pair at most once.
inline long long worst_list_items(int fnx, int fny, int fnz, int kchunk, int partial_min)
{
    return (long long)cap_tiles_x(fnx, partial_min) * cap_tiles_y(fny) * cap_nzchunk(fnz, kchunk);
}

// Entries of the overflow fix-up list in one step: each interior point at
// most once.
inline long long worst_fix_entries(int fnx, int fny, int fnz)
{
    return (long long)(fnx - 2) * (fny - 2) * (fnz - 2);
}

}  // namespace cupf
