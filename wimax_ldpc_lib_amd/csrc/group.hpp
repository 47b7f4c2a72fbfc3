// group.hpp -- one persistent decode kernel per lifting size z for mixed
// batches: the buckets of all six rates at that z are one work list.
//
// A mixed batch holds a few thousand frames per code; one persistent kernel
// per code then spends much of its time in its last round of tiles, where a
// tile of frames that do not converge runs all iterations while the rest of
// the chip waits.  Grouping the six codes of one z into one kernel (same n,
// so one LDS layout and one block size) gives each kernel six times the
// tiles.  The frames per workgroup are shared by the group (the value most of
// its rates were tuned to).
#pragma once

#include <hip/hip_runtime.h>

#include "qc_code.hpp"
#include "tdmp.hpp"
#include "tuned.hpp"

namespace wl {

struct GroupArgs {
    DecodeArgs a;                      // buffers, frame list, ragged offsets, iterations, early stop
    const uint32_t* buckets;           // [2 * 114] {first slot, count} per code, on the device
    unsigned long long* tile_counter;  // this group's tile counter (zeroed before the launch)
    unsigned enabled;                  // bit r: code (r, z) is decoded
    int frames_per_wg;                 // F, 1 .. GroupShape<Z>::fmax
};

// The group's launch shape.  Frames per workgroup are a launch argument (the
// tile body takes F at run time); the default is the value the mixed-batch
// autotuner measured for this z (csrc/group_tuned.inc), else the most common
// per-code tuned value over its rates, ties to the smaller block.
template <int Z>
struct GroupShape {
    // a rate's tuned F counts unless it was tuned to a wave-split shape (2
    // or 3), whose F is fixed by the shape (if all were, F = 2)
    static constexpr int rate_f(int r)
    {
        const int v = tuned_variant(r, Z);
        return (v == 2 || v == 3) ? 0 : (tuned_f(r, Z) > 0 ? tuned_f(r, Z) : 1);
    }
    static constexpr int pick_f()
    {
        int best = 0, best_n = 0;
        for (int r = 0; r < kNumRates; r++) {
            const int f = rate_f(r);
            if (f == 0) continue;
            int n = 0;
            for (int q = 0; q < kNumRates; q++) n += rate_f(q) == f ? 1 : 0;
            if (n > best_n || (n == best_n && f < best)) {
                best = f;
                best_n = n;
            }
        }
        return best > 0 ? best : 2;
    }
    static constexpr int heuristic_f() { return pick_f(); }
    static constexpr int F = group_tuned_f(Z) > 0 ? group_tuned_f(Z) : pick_f();
    // largest F: kMaxThreads lanes and 64 KiB of LDS
    static constexpr int max_f()
    {
        int f = kMaxThreads / Z;
        while (f > 1 && kLdsPad + (size_t)f * lds_frame_stride<24 * Z, VAR_LEAN>() * sizeof(float) > (size_t)kMaxLdsBytes) f--;
        return f;
    }
    static constexpr int fmax = max_f();
    // the register-lean shape for every rate: the hoisted-address shape
    // pushes the combined kernel past 168 VGPRs, and the two-halves shape
    // (six rates x one loop copy per wave) spills scalar registers and
    // measured slower here
    static constexpr int VAR = VAR_LEAN;
    static constexpr int threads(int f) { return ((f * Z + 63) / 64) * 64; }
    static constexpr size_t lds(int f) { return kLdsPad + (size_t)f * lds_frame_stride<24 * Z, VAR>() * sizeof(float); }
    template <int R>
    static constexpr int var()
    {
        return VAR;
    }
};

// Compiled for 4 waves per SIMD: 128 registers instead of the 160 the six
// rate bodies take unconstrained (3 waves), at the cost of some spills to
// scratch; on the 262144-frame mixed batch 4.43 against 4.64-4.74 ms.
template <int Z>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(4))) void tdmp_group_kernel(GroupArgs g)
{
    extern __shared__ __align__(16) unsigned char smem[];
    using S = GroupShape<Z>;
    const int F = g.frames_per_wg;
    constexpr int zi = index_of_z(Z);
    unsigned long long* slot = reinterpret_cast<unsigned long long*>(smem + kTileSlot);
    // Tiles of rate 0 first, then rate 1, ...: the counter only grows, so
    // every workgroup reaches the exit once the last tile is taken.  Nothing
    // is kept in LDS from one tile to the next: the tile index is written and
    // read between two barriers, and the tile counts per rate are re-read
    // from the device bucket table for every tile.  (A form that kept the
    // per-rate tile boundaries in the LDS pad for the whole kernel and
    // claimed each tile index one tile ahead decoded a few frames per
    // 262144-frame batch from the next code's bucket, at random; see the
    // full-size mixed test.)
    for (;;) {
        if (threadIdx.x == 0) *slot = atomicAdd(g.tile_counter, 1ull);
        __syncthreads();
        unsigned long long t = *slot;
        __syncthreads();
        int rate = -1;
        unsigned long long cnt = 0;
        for (int r = 0; r < kNumRates; r++) {
            const unsigned long long c = ((g.enabled >> r) & 1u) ? g.buckets[2 * (r * kNumZ + zi) + 1] : 0ull;
            const unsigned long long nt = (c + F - 1) / F;
            if (t < nt) {
                rate = r;
                cnt = c;
                break;
            }
            t -= nt;
        }
        if (rate < 0) break;   // uniform: t and the counts are the same in every thread
        const unsigned long long frame0 = t * F;
        const unsigned long long left = cnt - frame0;
        const int fe = left < (unsigned long long)F ? (int)left : F;
        static_for<0, kNumRates>([&](auto Rc) {
            constexpr int R = decltype(Rc)::value;
            if (rate != R) return;
            DecodeArgs a = g.a;
            a.bucket = g.buckets + 2 * (R * kNumZ + zi);
            a.frames_per_wg = F;
            tdmp_tile<R, Z, MODE_NMS, S::template var<R>(), true>(a, smem, frame0, fe);
        });
        __syncthreads();
    }
}

template <int Z>
hipError_t launch_group(const GroupArgs& g, int grid, hipStream_t stream)
{
    const int f = g.frames_per_wg;
    if (f < 1 || f > GroupShape<Z>::fmax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tdmp_group_kernel<Z>, dim3((unsigned)grid), dim3(GroupShape<Z>::threads(f)),
                       GroupShape<Z>::lds(f), stream, g);
    return hipGetLastError();
}

// workgroups of the group kernel resident per CU at F frames per workgroup (0 on error)
template <int Z>
int occupancy_group(int f)
{
    int nb = 0;
    if (f < 1 || f > GroupShape<Z>::fmax) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tdmp_group_kernel<Z>, GroupShape<Z>::threads(f),
                                                     GroupShape<Z>::lds(f)) != hipSuccess)
        return 0;
    return nb;
}

using GroupLaunch = hipError_t (*)(const GroupArgs&, int, hipStream_t);
using GroupOccupancy = int (*)(int);

struct GroupEntry {
    GroupLaunch launch;
    GroupOccupancy occ;
    int frames_per_wg;   // default (tuned) F
    int fmax;
};

// defined in decode_g<k>.hip, one per lifting size
#define WL_DECLARE_GROUP(Z) const GroupEntry* group_entry_z##Z();
WL_DECLARE_GROUP(24) WL_DECLARE_GROUP(28) WL_DECLARE_GROUP(32) WL_DECLARE_GROUP(36) WL_DECLARE_GROUP(40)
WL_DECLARE_GROUP(44) WL_DECLARE_GROUP(48) WL_DECLARE_GROUP(52) WL_DECLARE_GROUP(56) WL_DECLARE_GROUP(60)
WL_DECLARE_GROUP(64) WL_DECLARE_GROUP(68) WL_DECLARE_GROUP(72) WL_DECLARE_GROUP(76) WL_DECLARE_GROUP(80)
WL_DECLARE_GROUP(84) WL_DECLARE_GROUP(88) WL_DECLARE_GROUP(92) WL_DECLARE_GROUP(96)
#undef WL_DECLARE_GROUP

inline const GroupEntry* group_entry(int z)
{
    switch (z) {
#define WL_GROUP_CASE(Z) \
    case Z: return group_entry_z##Z();
        WL_GROUP_CASE(24) WL_GROUP_CASE(28) WL_GROUP_CASE(32) WL_GROUP_CASE(36) WL_GROUP_CASE(40)
        WL_GROUP_CASE(44) WL_GROUP_CASE(48) WL_GROUP_CASE(52) WL_GROUP_CASE(56) WL_GROUP_CASE(60)
        WL_GROUP_CASE(64) WL_GROUP_CASE(68) WL_GROUP_CASE(72) WL_GROUP_CASE(76) WL_GROUP_CASE(80)
        WL_GROUP_CASE(84) WL_GROUP_CASE(88) WL_GROUP_CASE(92) WL_GROUP_CASE(96)
#undef WL_GROUP_CASE
    default: return nullptr;
    }
}

}  // namespace wl

#define WL_DEFINE_GROUP(Z)                                                                  \
    namespace wl {                                                                          \
    const GroupEntry* group_entry_z##Z()                                                    \
    {                                                                                       \
        static const GroupEntry e{&launch_group<Z>, &occupancy_group<Z>, GroupShape<Z>::F, GroupShape<Z>::fmax}; \
        return &e;                                                                          \
    }                                                                                       \
    }
