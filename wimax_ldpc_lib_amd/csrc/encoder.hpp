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

hipError_t launch_encode(const EncodeDesc& d, const uint8_t* info, uint8_t* cw,
                         unsigned long long nframes, hipStream_t s);
hipError_t launch_channel(const EncodeDesc& d, unsigned long long seed, unsigned long long frame_offset,
                          unsigned long long nframes, float eb_lo, float eb_hi, float* llr,
                          uint8_t* bits, float* ebn0, hipStream_t s,
                          const unsigned long long* step_dev = nullptr);

}  // namespace wl
