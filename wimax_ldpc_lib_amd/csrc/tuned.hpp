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

// frames per workgroup of the mixed-batch group kernel of lifting size z,
// measured by the autotuner on mixed batches (group_tuned.inc: {z, F}); 0
// when the table has no entry
struct GroupTuned {
    int z, f;
};

constexpr GroupTuned kGroupTuned[] = {
#include "group_tuned.inc"
    {0, 0}};

constexpr int group_tuned_f(int z)
{
    for (const GroupTuned& g : kGroupTuned)
        if (g.z == z) return g.f;
    return 0;
}

}  // namespace wl
