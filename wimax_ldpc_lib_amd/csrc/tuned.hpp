// tuned.hpp -- the autotuned launch shapes as a compile-time table.
//
// tuned_table.inc is written by wimax_ldpc_lib_amd/autotune.py from
// measurements on MI355X: {rate, z, frames_per_wg, variant} per code.
#pragma once

#include "qc_code.hpp"

namespace wl {

struct TunedShape {
    int rate, z, f, variant;
};

constexpr TunedShape kTunedShapes[] = {
#include "tuned_table.inc"
    {-1, 0, 0, 0}};

// frames per workgroup / code-shape variant (0, 1) of code (rate, z); 0 / -1
// when the table has no entry
constexpr int tuned_f(int rate, int z)
{
    for (const TunedShape& t : kTunedShapes)
        if (t.rate == rate && t.z == z) return t.f;
    return 0;
}

constexpr int tuned_variant(int rate, int z)
{
    for (const TunedShape& t : kTunedShapes)
        if (t.rate == rate && t.z == z) return t.variant;
    return -1;
}

}  // namespace wl
