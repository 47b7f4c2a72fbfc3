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
    // std_form = 1: the general dual-diagonal back-substitution, with the
    // first parity column's shifts h_top (base rows 0 and mb-1) and h_mid
    // (base row x); 0: the reference's parity prediction and correction
    int std_form, h_top, h_mid;
    // compiled = 1: an 802.16e code (rate, z) whose encoder kernel is
    // instantiated with its structure as constants; 0: run-time code
    int compiled, rate;
    int8_t deg[kMaxLayers];
    int8_t blk[kMaxLayers][kMaxDeg];
    int8_t sh[kMaxLayers][kMaxDeg];
};

// encoder descriptor of (rate, z); frames_per_wg = 256 / z
EncodeDesc make_encode_desc_for(int rate, int z);
// the general back-substitution form of the same code (valid for 3/4B too);
// false if the parity part does not have the dual-diagonal shape
bool make_standard_encode_desc(int rate, int z, EncodeDesc* out);

hipError_t launch_encode(const EncodeDesc& d, const uint8_t* info, uint8_t* cw,
                         unsigned long long nframes, hipStream_t s);
hipError_t launch_channel(const EncodeDesc& d, unsigned long long seed, unsigned long long frame_offset,
                          unsigned long long nframes, float eb_lo, float eb_hi, float* llr,
                          uint8_t* bits, float* ebn0, hipStream_t s,
                          const unsigned long long* step_dev = nullptr);

}  // namespace wl
