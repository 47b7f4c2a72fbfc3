// encoder.hip -- dual-diagonal QC-LDPC encoder and synthetic channel kernels.
//
// encode_kernel restates ldpc_encoder::encode_data (lib/ldpc_encoder.cc:50-105),
// the Lin-Wei-Ku parity prediction and correction, block-parallel: lane t of
// a frame owns position t of every z-block, so
//   V_q[t]  = XOR over info slots (l = q, blk < kb) of u[blk*z + (t + s) mod z]  (:29-48)
//   P_q     = V_q                       for the last block q = mb-1             (:61-64)
//   P_q     = P_{q-1} ^ V_{q-1}         for q = 1 .. X  (P_0 = 0)              (:67-70)
//   P_q     = V_q ^ P_{q+1}             for q = mb-2 down to X+1               (:73-76)
//   P_0     = V_X ^ P_{X+1} ^ P_X                                              (:79-88)
//   c[k + q z + t] = P_q[t] ^ P_0[(t + D) mod z] for q >= 1, P_0[t] for q = 0 (:93-103)
// and the systematic part is the info bytes as given (:90).  All integer XOR,
// bit-exact with the reference.
//
// channel_kernel fuses a seeded source, the encoder and a BPSK/AWGN channel,
// the GPU counterpart of lib/BER_runner.cc:162-191 (which draws its bits from
// an LFSR and its noise from std::mt19937; those streams are not reproduced).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "channel_dev.hpp"
#include "encoder.hpp"

namespace wl {

__global__ __launch_bounds__(256) void encode_kernel(EncodeDesc d, const uint8_t* __restrict__ info,
                                                     uint8_t* __restrict__ cw,
                                                     unsigned long long nframes)
{
    extern __shared__ uint8_t esm[];
    const int z = d.z, k = d.k, n = d.n, F = d.frames_per_wg;
    const unsigned long long frame0 = (unsigned long long)blockIdx.x * F;
    const unsigned long long left = nframes - frame0;
    const int Feff = left < (unsigned long long)F ? (int)left : F;
    const int tid = threadIdx.x;
    const int fs = tid / z, t = tid - fs * z;
    const bool active = fs < Feff;
    uint8_t* u = esm;                         // [F][k]
    uint8_t* p0 = esm + (size_t)F * k;        // [F][z]
    const uint8_t* src = info + frame0 * k;
    uint8_t* dst = cw + frame0 * n;
    for (int i = tid; i < Feff * k; i += blockDim.x) {
        const uint8_t b = src[i];
        u[i] = b != 0;
        const int f = i / k, j = i - f * k;
        dst[(size_t)f * n + j] = b;           // systematic part copied as is
    }
    __syncthreads();
    const int fsa = active ? fs : 0;
    encode_frame(d, u + (size_t)fsa * k, p0 + (size_t)fsa * z, t, active,
                 [&](int i, uint8_t bit) { dst[(size_t)fs * n + k + i] = bit; });
}

__global__ __launch_bounds__(256) void channel_kernel(EncodeDesc d, unsigned long long seed,
                                                      unsigned long long frame_offset,
                                                      unsigned long long nframes, float eb_lo,
                                                      float eb_hi, float* __restrict__ llr,
                                                      uint8_t* __restrict__ bits,
                                                      float* __restrict__ ebn0_out,
                                                      const unsigned long long* __restrict__ step_dev)
{
    extern __shared__ uint8_t esm[];
    const int z = d.z, k = d.k, n = d.n, F = d.frames_per_wg;
    // graph replays advance through fresh frames: absolute frame = offset + step * nframes
    if (step_dev) frame_offset += *step_dev * nframes;
    const unsigned long long frame0 = (unsigned long long)blockIdx.x * F;
    const unsigned long long left = nframes - frame0;
    const int Feff = left < (unsigned long long)F ? (int)left : F;
    const int tid = threadIdx.x;
    const int fs = tid / z, t = tid - fs * z;
    const bool active = fs < Feff;
    uint8_t* c = esm;                         // [F][n] codeword bits
    uint8_t* p0 = esm + (size_t)F * n;        // [F][z]
    const int nb = (k + 127) / 128;
    for (int i = tid; i < Feff * nb; i += blockDim.x) {
        const int f = i / nb, blk = i - f * nb;
        uint32_t w[4];
        info_block(seed, frame_offset + frame0 + f, blk, w);
        for (int q = 0; q < 128; q++) {
            const int j = blk * 128 + q;
            if (j < k) c[(size_t)f * n + j] = (w[q >> 5] >> (q & 31)) & 1u;
        }
    }
    __syncthreads();
    const int fsa = active ? fs : 0;
    encode_frame(d, c + (size_t)fsa * n, p0 + (size_t)fsa * z, t, active,
                 [&](int i, uint8_t bit) { c[(size_t)fs * n + k + i] = bit; });
    __syncthreads();
    for (int f = 0; f < Feff; f++) {
        const unsigned long long g = frame_offset + frame0 + f;
        const float eb = frame_ebn0(seed, g, eb_lo, eb_hi);
        const double s2 = frame_s2(d, eb);
        const double sig = sqrt(s2);
        if (tid == 0 && ebn0_out) ebn0_out[frame0 + f] = eb;
        for (int v2 = tid; v2 < n / 2; v2 += blockDim.x) {
            const uint8_t b0 = c[(size_t)f * n + 2 * v2], b1 = c[(size_t)f * n + 2 * v2 + 1];
            if (llr) reinterpret_cast<float2*>(llr + (frame0 + f) * n)[v2] = noisy_pair(seed, g, v2, b0, b1, s2, sig);
            if (bits) {
                bits[(frame0 + f) * n + 2 * v2] = b0;
                bits[(frame0 + f) * n + 2 * v2 + 1] = b1;
            }
        }
    }
}

EncodeDesc make_encode_desc_for(int rate, int z)
{
    const LayerTable lt = make_layer_table(rate);
    const wl_base_info& bi = wl_base_table[rate];
    EncodeDesc d{};
    d.z = z;
    d.n = 24 * z;
    d.m = lt.nl * z;
    d.k = d.n - d.m;
    d.mb = lt.nl;
    d.kb = 24 - lt.nl;
    d.x = bi.x;
    d.d = lift(rate, bi.d0, z);
    d.frames_per_wg = 256 / z > 0 ? 256 / z : 1;
    for (int l = 0; l < lt.nl; l++) {
        d.deg[l] = (int8_t)lt.deg[l];
        for (int j = 0; j < lt.deg[l]; j++) {
            d.blk[l][j] = (int8_t)lt.blk[l][j];
            d.sh[l][j] = (int8_t)lift(rate, lt.p96[l][j], z);
        }
    }
    return d;
}

bool make_standard_encode_desc(int rate, int z, EncodeDesc* out)
{
    EncodeDesc d = make_encode_desc_for(rate, z);
    const LayerTable lt = make_layer_table(rate);
    // the first parity column (block kb): shifts at base rows 0, x and mb-1
    // only, equal at 0 and mb-1; block kb + j (j >= 1): shift-0 blocks at
    // rows j-1 and j only
    int h[kMaxLayers];
    for (int l = 0; l < lt.nl; l++) {
        h[l] = -1;
        bool stair_ok = true;
        for (int j = 0; j < lt.deg[l]; j++) {
            const int b = lt.blk[l][j];
            if (b == d.kb) h[l] = lift(rate, lt.p96[l][j], z);
            if (b > d.kb && (lt.p96[l][j] != 0 || (b - d.kb != l && b - d.kb != l + 1))) stair_ok = false;
        }
        if (!stair_ok) return false;
    }
    for (int l = 1; l < lt.nl - 1; l++)
        if (l != d.x && h[l] >= 0) return false;
    if (h[0] < 0 || h[d.x] < 0 || h[0] != h[lt.nl - 1]) return false;
    d.std_form = 1;
    d.h_top = h[0];
    d.h_mid = h[d.x];
    *out = d;
    return true;
}

hipError_t launch_encode(const EncodeDesc& d, const uint8_t* info, uint8_t* cw,
                         unsigned long long nframes, hipStream_t s)
{
    if (nframes == 0) return hipSuccess;
    const int F = d.frames_per_wg;
    const unsigned long long grid = (nframes + F - 1) / F;
    const size_t lds = (size_t)F * (d.k + d.z);
    hipLaunchKernelGGL(encode_kernel, dim3((unsigned)grid), dim3(((F * d.z + 63) / 64) * 64), lds, s,
                       d, info, cw, nframes);
    return hipGetLastError();
}

hipError_t launch_channel(const EncodeDesc& d, unsigned long long seed, unsigned long long frame_offset,
                          unsigned long long nframes, float eb_lo, float eb_hi, float* llr,
                          uint8_t* bits, float* ebn0, hipStream_t s, const unsigned long long* step_dev)
{
    if (nframes == 0) return hipSuccess;
    const int F = d.frames_per_wg;
    const unsigned long long grid = (nframes + F - 1) / F;
    const size_t lds = (size_t)F * (d.n + d.z);
    hipLaunchKernelGGL(channel_kernel, dim3((unsigned)grid), dim3(((F * d.z + 63) / 64) * 64), lds, s,
                       d, seed, frame_offset, nframes, eb_lo, eb_hi, llr, bits, ebn0, step_dev);
    return hipGetLastError();
}

}  // namespace wl
