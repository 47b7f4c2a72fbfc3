// generic.hpp -- layered normalised min-sum decoding of an arbitrary
// quasi-cyclic LDPC code described at run time (base matrix of circulant
// shifts), for codes beyond the compiled 802.16e tables: codes read from an
// alist file, 802.11n-style base matrices, other lifting sizes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wl {

constexpr int kGenMaxLayers = 64;   // base rows
constexpr int kGenMaxDeg = 32;      // nonzero blocks per base row
constexpr int kGenMaxZ = 512;       // lifting size (one lane per row of a layer)
constexpr int kGenLdsPad = 64;      // per-frame counters
constexpr int kGenMaxLds = 65536;

// Device-resident code description: layer l has deg[l] nonzero blocks, slot j
// in block column blk[l][j] (ascending) with circulant shift sh[l][j]; row r
// of layer l touches variable blk*z + (r + sh) mod z.  Messages of layer l
// start at float moff[l] of the message area, slot-major, z per slot.
struct GenDesc {
    int z, nb, mb, n, m, edges;
    int deg[kGenMaxLayers];
    int moff[kGenMaxLayers];
    int16_t blk[kGenMaxLayers][kGenMaxDeg];
    int16_t sh[kGenMaxLayers][kGenMaxDeg];
};

struct GenArgs {
    const GenDesc* desc;   // device copy
    const float* llr;      // [nframes][n], reference sign convention
    float* post;           // [nframes][n] (nullable; may alias llr)
    uint8_t* hard;         // [nframes][n/8] packed, bit v%8 of byte v/8 (nullable; n % 8 == 0)
    uint8_t* ret;          // [nframes] (nullable)
    uint8_t* iters;        // [nframes] (nullable)
    uint16_t* synd;        // [nframes] (nullable)
    unsigned long long nframes;
    int max_iter;
    int early_stop;
    int n, edges, z;       // host copies, for the launch shape
};

// LDS bytes one frame needs (counters, posteriors, messages)
inline size_t gen_lds_bytes(int n, int edges) { return kGenLdsPad + ((size_t)n + (size_t)edges) * sizeof(float); }

hipError_t launch_generic(const GenArgs& a, hipStream_t stream);

}  // namespace wl
