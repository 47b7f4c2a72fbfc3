// encoder.hpp -- run-time code descriptor and launchers of the encoder and
// channel kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qc_code.hpp"

namespace wl {

struct EncodeDesc {
    int z, n, k, m, mb, kb, x, d;
    int frames_per_wg;
    int8_t deg[kMaxLayers];
    int8_t blk[kMaxLayers][kMaxDeg];
    int8_t sh[kMaxLayers][kMaxDeg];
};

// encoder descriptor of (rate, z); frames_per_wg = 256 / z
EncodeDesc make_encode_desc_for(int rate, int z);

hipError_t launch_encode(const EncodeDesc& d, const uint8_t* info, uint8_t* cw,
                         unsigned long long nframes, hipStream_t s);
hipError_t launch_channel(const EncodeDesc& d, unsigned long long seed, unsigned long long frame_offset,
                          unsigned long long nframes, float eb_lo, float eb_hi, float* llr,
                          uint8_t* bits, float* ebn0, hipStream_t s,
                          const unsigned long long* step_dev = nullptr);

}  // namespace wl
