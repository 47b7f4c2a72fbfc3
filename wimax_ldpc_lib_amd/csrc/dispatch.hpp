// dispatch.hpp -- run-time (rate, z) -> compiled kernel tables.
#pragma once

#include <hip/hip_runtime.h>

#include "qc_code.hpp"
#include "tdmp.hpp"

namespace wl {

using DecodeLaunch = hipError_t (*)(const DecodeArgs&, hipStream_t);
using DecodeOccupancy = int (*)(int frames_per_wg);

// code-shape variants of the NMS kernel (tdmp.hpp VAR_*): 0 hoisted row
// addresses, 1 VAR_LEAN, 2 VAR_LEAN_W2 (two frames per workgroup, z % 32 == 0
// only), 3 VAR_LEAN_W4 (four frames per workgroup, z = 32, 48, 64 only: one
// loop copy per wave, at most four), 4 the hoisted shape compiled for 4 waves
// per SIMD (contiguous batches; frame lists and persistent launches take
// variant 0); null where a shape does not exist
constexpr int kNumVariants = 5;
constexpr int kVariantW2 = 2;
constexpr int kVariantW4 = 3;

// frames per workgroup a variant requires (0: any)
constexpr int variant_frames(int variant)
{
    return variant == kVariantW2 ? 2 : (variant == kVariantW4 ? 4 : 0);
}

struct DecodeEntry {
    DecodeLaunch nms[kNumVariants];  // normalised min-sum, the reference's arithmetic
    DecodeLaunch learned;            // learned per-block offset min-sum (FP32)
    int fmax;                        // largest frames-per-workgroup the kernel accepts
    DecodeOccupancy occ[kNumVariants];
    DecodeOccupancy occ_learned;
    DecodeLaunch learned_fb;         // learned + in-kernel normalised min-sum re-queue
                                     // (persistent only; built for z = 24 and 96, else null)
    DecodeOccupancy occ_learned_fb;
    DecodeLaunch alt;                // offset / plain min-sum (DecodeArgs::rule)
};

// learned + re-queue: the two-halves shape where the decoder runs two frames
// per workgroup and z % 32 == 0 (the shape tuned for every z = 96 code; both
// the learned and the re-queued tiles use it), the register-lean shape
// otherwise
template <int RATE, int Z>
hipError_t launch_learned_fb(const DecodeArgs& a, hipStream_t stream)
{
    if constexpr (Z % 32 == 0)
        if (a.frames_per_wg == 2) return launch_tdmp<RATE, Z, MODE_LEARNED_FB, VAR_LEAN_W2>(a, stream);
    return launch_tdmp<RATE, Z, MODE_LEARNED_FB, VAR_LEAN>(a, stream);
}

template <int RATE, int Z>
int occupancy_learned_fb(int frames_per_wg)
{
    if constexpr (Z % 32 == 0)
        if (frames_per_wg == 2) return occupancy_tdmp<RATE, Z, MODE_LEARNED_FB, VAR_LEAN_W2, true>(frames_per_wg);
    return occupancy_tdmp<RATE, Z, MODE_LEARNED_FB, VAR_LEAN, true>(frames_per_wg);
}

template <int RATE, int Z>
constexpr DecodeLaunch fb_launcher()
{
    if constexpr (Z == 24 || Z == 96) return &launch_learned_fb<RATE, Z>;
    else return nullptr;
}

template <int RATE, int Z>
constexpr DecodeLaunch w2_launcher()
{
    if constexpr (Z % 32 == 0) return &launch_tdmp<RATE, Z, MODE_NMS, VAR_LEAN_W2>;
    else return nullptr;
}

template <int RATE, int Z>
constexpr DecodeOccupancy w2_occupancy()
{
    if constexpr (Z % 32 == 0) return &occupancy_tdmp<RATE, Z, MODE_NMS, VAR_LEAN_W2, true>;
    else return nullptr;
}

template <int RATE, int Z>
constexpr DecodeLaunch w4_launcher()
{
    if constexpr (Z % 16 == 0 && Z <= 64) return &launch_tdmp<RATE, Z, MODE_NMS, VAR_LEAN_W4>;
    else return nullptr;
}

template <int RATE, int Z>
constexpr DecodeOccupancy w4_occupancy()
{
    if constexpr (Z % 16 == 0 && Z <= 64) return &occupancy_tdmp<RATE, Z, MODE_NMS, VAR_LEAN_W4, true>;
    else return nullptr;
}

// learned inference: the two-halves shape where the decoder runs two frames
// per workgroup and z % 32 == 0 (as the normalised min-sum kernel tuned
// there), the hoisted-address shape otherwise
template <int RATE, int Z>
hipError_t launch_learned(const DecodeArgs& a, hipStream_t stream)
{
    if constexpr (Z % 32 == 0)
        if (a.frames_per_wg == 2) return launch_tdmp<RATE, Z, MODE_LEARNED, VAR_LEAN_W2>(a, stream);
    return launch_tdmp<RATE, Z, MODE_LEARNED, 0>(a, stream);
}

// variant 4: contiguous batches in the hoisted shape at 4 waves per SIMD,
// everything else in variant 0's extended kernel
template <int RATE, int Z>
hipError_t launch_hoist4(const DecodeArgs& a, hipStream_t stream)
{
    if (a.frame_idx || a.bucket || a.frame_off || a.tile_counter) return launch_tdmp<RATE, Z, MODE_NMS, 0>(a, stream);
    return launch_tdmp_plain<RATE, Z, MODE_NMS, VAR_HOIST4>(a, stream);
}

// offset / plain min-sum: the two-halves shape where the decoder runs two
// frames per workgroup and z % 32 == 0 (as the other rules there), the
// register-lean shape otherwise
template <int RATE, int Z>
hipError_t launch_alt(const DecodeArgs& a, hipStream_t stream)
{
    if constexpr (Z % 32 == 0)
        if (a.frames_per_wg == 2) return launch_tdmp<RATE, Z, MODE_ALT, VAR_LEAN_W2>(a, stream);
    return launch_tdmp<RATE, Z, MODE_ALT, VAR_LEAN>(a, stream);
}

template <int RATE, int Z>
constexpr DecodeOccupancy fb_occupancy()
{
    if constexpr (Z == 24 || Z == 96) return &occupancy_learned_fb<RATE, Z>;
    else return nullptr;
}

// Each rate's kNumZ entries come in two parts, compiled in two units so
// that a parallel build spreads the kernels: lifting sizes 24..60 (z index
// 0..kZSplit-1) in decode_r<rate>_lo.hip, 64..96 in decode_r<rate>_hi.hip.
constexpr int kZSplit = 10;
const DecodeEntry* decode_table_r0_lo();
const DecodeEntry* decode_table_r1_lo();
const DecodeEntry* decode_table_r2_lo();
const DecodeEntry* decode_table_r3_lo();
const DecodeEntry* decode_table_r4_lo();
const DecodeEntry* decode_table_r5_lo();
const DecodeEntry* decode_table_r0_hi();
const DecodeEntry* decode_table_r1_hi();
const DecodeEntry* decode_table_r2_hi();
const DecodeEntry* decode_table_r3_hi();
const DecodeEntry* decode_table_r4_hi();
const DecodeEntry* decode_table_r5_hi();

inline const DecodeEntry* decode_entry(int rate, int z)
{
    if (rate < 0 || rate >= kNumRates || !valid_z(z)) return nullptr;
    const int zi = index_of_z(z);
    const bool hi = zi >= kZSplit;
    const DecodeEntry* t = nullptr;
    switch (rate) {
    case 0: t = hi ? decode_table_r0_hi() : decode_table_r0_lo(); break;
    case 1: t = hi ? decode_table_r1_hi() : decode_table_r1_lo(); break;
    case 2: t = hi ? decode_table_r2_hi() : decode_table_r2_lo(); break;
    case 3: t = hi ? decode_table_r3_hi() : decode_table_r3_lo(); break;
    case 4: t = hi ? decode_table_r4_hi() : decode_table_r4_lo(); break;
    default: t = hi ? decode_table_r5_hi() : decode_table_r5_lo(); break;
    }
    return &t[hi ? zi - kZSplit : zi];
}

// entries of z index LO + i for i in I...
template <int RATE, int LO, int... I>
constexpr auto make_decode_table(std::integer_sequence<int, I...>)
{
    struct Arr { DecodeEntry e[sizeof...(I)]; };
    return Arr{{DecodeEntry{{&launch_tdmp<RATE, z_of_index(LO + I), MODE_NMS, 0>,
                             &launch_tdmp<RATE, z_of_index(LO + I), MODE_NMS, VAR_LEAN>,
                             w2_launcher<RATE, z_of_index(LO + I)>(), w4_launcher<RATE, z_of_index(LO + I)>(),
                             &launch_hoist4<RATE, z_of_index(LO + I)>},
                            &launch_learned<RATE, z_of_index(LO + I)>,
                            max_frames_per_wg<RATE, z_of_index(LO + I)>(),
                            {&occupancy_tdmp<RATE, z_of_index(LO + I), MODE_NMS, 0, true>,
                             &occupancy_tdmp<RATE, z_of_index(LO + I), MODE_NMS, VAR_LEAN, true>,
                             w2_occupancy<RATE, z_of_index(LO + I)>(), w4_occupancy<RATE, z_of_index(LO + I)>(),
                             &occupancy_tdmp<RATE, z_of_index(LO + I), MODE_NMS, 0, true>},
                            &occupancy_tdmp<RATE, z_of_index(LO + I), MODE_LEARNED, 0, true>,
                            fb_launcher<RATE, z_of_index(LO + I)>(),
                            fb_occupancy<RATE, z_of_index(LO + I)>(),
                            &launch_alt<RATE, z_of_index(LO + I)>}...}};
}

}  // namespace wl

// PART lo: z index 0 .. kZSplit-1; PART hi: kZSplit .. kNumZ-1
#define WL_DEFINE_DECODE_TABLE_lo(RATE)                                                     \
    namespace wl {                                                                          \
    const DecodeEntry* decode_table_r##RATE##_lo()                                          \
    {                                                                                       \
        static const auto tab = make_decode_table<RATE, 0>(std::make_integer_sequence<int, kZSplit>{}); \
        return tab.e;                                                                       \
    }                                                                                       \
    }
#define WL_DEFINE_DECODE_TABLE_hi(RATE)                                                     \
    namespace wl {                                                                          \
    const DecodeEntry* decode_table_r##RATE##_hi()                                          \
    {                                                                                       \
        static const auto tab =                                                             \
            make_decode_table<RATE, kZSplit>(std::make_integer_sequence<int, kNumZ - kZSplit>{}); \
        return tab.e;                                                                       \
    }                                                                                       \
    }
