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

// One table of kNumZ entries per rate, defined in decode_r<rate>.hip.
const DecodeEntry* decode_table_r0();
const DecodeEntry* decode_table_r1();
const DecodeEntry* decode_table_r2();
const DecodeEntry* decode_table_r3();
const DecodeEntry* decode_table_r4();
const DecodeEntry* decode_table_r5();

inline const DecodeEntry* decode_entry(int rate, int z)
{
    if (rate < 0 || rate >= kNumRates || !valid_z(z)) return nullptr;
    const DecodeEntry* t = nullptr;
    switch (rate) {
    case 0: t = decode_table_r0(); break;
    case 1: t = decode_table_r1(); break;
    case 2: t = decode_table_r2(); break;
    case 3: t = decode_table_r3(); break;
    case 4: t = decode_table_r4(); break;
    default: t = decode_table_r5(); break;
    }
    return &t[index_of_z(z)];
}

template <int RATE, int... ZI>
constexpr auto make_decode_table(std::integer_sequence<int, ZI...>)
{
    struct Arr { DecodeEntry e[sizeof...(ZI)]; };
    return Arr{{DecodeEntry{{&launch_tdmp<RATE, z_of_index(ZI), MODE_NMS, 0>,
                             &launch_tdmp<RATE, z_of_index(ZI), MODE_NMS, VAR_LEAN>,
                             w2_launcher<RATE, z_of_index(ZI)>(), w4_launcher<RATE, z_of_index(ZI)>(),
                             &launch_hoist4<RATE, z_of_index(ZI)>},
                            &launch_learned<RATE, z_of_index(ZI)>,
                            max_frames_per_wg<RATE, z_of_index(ZI)>(),
                            {&occupancy_tdmp<RATE, z_of_index(ZI), MODE_NMS, 0, true>,
                             &occupancy_tdmp<RATE, z_of_index(ZI), MODE_NMS, VAR_LEAN, true>,
                             w2_occupancy<RATE, z_of_index(ZI)>(), w4_occupancy<RATE, z_of_index(ZI)>(),
                             &occupancy_tdmp<RATE, z_of_index(ZI), MODE_NMS, 0, true>},
                            &occupancy_tdmp<RATE, z_of_index(ZI), MODE_LEARNED, 0, true>,
                            fb_launcher<RATE, z_of_index(ZI)>(),
                            fb_occupancy<RATE, z_of_index(ZI)>(),
                            &launch_alt<RATE, z_of_index(ZI)>}...}};
}

}  // namespace wl

#define WL_DEFINE_DECODE_TABLE(RATE)                                                        \
    namespace wl {                                                                          \
    const DecodeEntry* decode_table_r##RATE()                                               \
    {                                                                                       \
        static const auto tab = make_decode_table<RATE>(std::make_integer_sequence<int, kNumZ>{}); \
        return tab.e;                                                                       \
    }                                                                                       \
    }
