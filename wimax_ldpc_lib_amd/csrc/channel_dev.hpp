// channel_dev.hpp -- device bodies shared by the encoder, channel and mixed
// kernels: the dual-diagonal encoder of one frame and the seeded source and
// BPSK/AWGN channel (included by each kernel file; no relocatable device code).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "encoder.hpp"
#include "philox.hpp"

namespace wl {

// lambda_q[t] of lane t: the XOR over the info blocks of check row q t of
// u[blk z + (t + s) mod z], u the frame's 0/1 info bits in LDS
// (lib/ldpc_encoder.cc:29-48).  Run-time descriptor.
__device__ __forceinline__ void encode_lambda(const EncodeDesc& d, const uint8_t* u, int t, bool active,
                                              uint8_t (&V)[kMaxLayers])
{
    const int z = d.z;
#pragma unroll
    for (int q = 0; q < kMaxLayers; q++) {
        uint8_t v = 0;
        if (q < d.mb && active) {
            for (int j = 0; j < d.deg[q]; j++) {
                const int b = d.blk[q][j];
                if (b >= d.kb) break;            // slots ascend by column: info blocks first
                int p = t + d.sh[q][j];
                p = p >= z ? p - z : p;
                v ^= u[b * z + p];
            }
        }
        V[q] = v;
    }
}

// The same for the compiled code (RATE, Z): every slot's block and shift are
// constants, so all of a lane's LDS reads can be in flight at once.
template <int RATE, int Z>
__device__ __forceinline__ void encode_lambda_ct(const uint8_t* u, int t, bool active, uint8_t (&V)[kMaxLayers])
{
    using CD = Code<RATE, Z>;
    constexpr int KB = 24 - CD::NL;
#pragma unroll
    for (int q = 0; q < kMaxLayers; q++) V[q] = 0;
    if (!active) return;
    static_for<0, CD::NL>([&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        uint8_t v = 0;
        static_for<0, CD::deg(L)>([&](auto Jc) {
            constexpr int J = decltype(Jc)::value;
            constexpr int b = CD::blk(L, J), s = CD::sh(L, J);
            if constexpr (b < KB) {
                const int p = t >= Z - s ? t + s - Z : t + s;
                v ^= u[b * Z + p];
            }
        });
        V[L] = v;
    });
}

// Parity bits of the frame from its lambda blocks, lane t < z; P_0 scratch at
// p0 (z words).  Writes bit i of the parity part through put(i, bit),
// i in [0, m).  All threads of the workgroup call it (one barrier).  T is
// uint8_t (one frame, a 0/1 byte per bit) or a bit-sliced word (bit f of
// every word belongs to frame f of a tile): the steps are XORs only, so
// each bit position of T runs the same encoder.
template <class T, class Put>
__device__ __forceinline__ void encode_parity(const EncodeDesc& d, const T (&V)[kMaxLayers], T* p0,
                                              int t, bool active, Put&& put)
{
    const int z = d.z;
    T P[kMaxLayers];
#pragma unroll
    for (int q = 0; q < kMaxLayers; q++) P[q] = 0;
    const int mb = d.mb, X = d.x;
    if (d.std_form) {
        // Summing all check rows cancels every staircase block and the two
        // h_top blocks of the first parity column: P^h_mid p0 = XOR of the
        // lambda_q, so p0[t] = S[(t - h_mid) mod z].  Then row 0 gives
        // p_1 = lambda_0 + P^h_top p0 and row q (1 <= q <= mb-2)
        // p_{q+1} = lambda_q + p_q (+ P^h_mid p0 = S at q = x).
        T S = 0;
#pragma unroll
        for (int q = 0; q < kMaxLayers; q++)
            if (q < mb) S ^= V[q];
        if (active) p0[t] = S;
        __syncthreads();
        if (active) {
            int a0 = t - d.h_mid;
            a0 += a0 < 0 ? z : 0;
            int a1 = t + d.h_top - d.h_mid;
            a1 += a1 < 0 ? z : 0;
            a1 -= a1 >= z ? z : 0;
            put(t, p0[a0]);
            T p = V[0] ^ p0[a1];
            put(z + t, p);
#pragma unroll
            for (int q = 1; q < kMaxLayers - 1; q++)
                if (q <= mb - 2) {
                    p = V[q] ^ p ^ (q == X ? S : (T)0);
                    put((q + 1) * z + t, p);
                }
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < kMaxLayers; q++)
        if (q == mb - 1) P[q] = V[q];
#pragma unroll
    for (int q = 1; q < kMaxLayers; q++)
        if (q <= X) P[q] = P[q - 1] ^ V[q - 1];
#pragma unroll
    for (int q = kMaxLayers - 2; q >= 1; q--)
        if (q <= mb - 2 && q >= X + 1) P[q] = V[q] ^ P[q + 1];
    T vx = 0, px1 = 0, px = 0;
#pragma unroll
    for (int q = 0; q < kMaxLayers; q++) {
        if (q == X) { vx = V[q]; px = P[q]; }
        if (q == X + 1) px1 = P[q];
    }
    const T pz = vx ^ px1 ^ px;
    if (active) p0[t] = pz;
    __syncthreads();
    if (active) {
        int tr = t + d.d;
        tr %= z;
        const T corr = p0[tr];
        put(t, pz);
#pragma unroll
        for (int q = 1; q < kMaxLayers; q++)
            if (q < mb) put(q * z + t, (T)(P[q] ^ corr));
    }
}

// Shared encoder body: info bytes (already 0/1) of frame slot fs in LDS at
// u, P_0 scratch at p0 (z bytes); lane t < z.  Writes the parity bits through
// put(i, bit) for i in [0, m).
template <class Put>
__device__ __forceinline__ void encode_frame(const EncodeDesc& d, const uint8_t* u, uint8_t* p0,
                                             int t, bool active, Put&& put)
{
    uint8_t V[kMaxLayers];
    encode_lambda(d, u, t, active, V);
    encode_parity(d, V, p0, t, active, put);
}

// Source and channel of absolute frame g (Philox4x32-10 keyed by seed):
// info bit j from counter (g, j / 128, 0), word (j % 128) / 32, bit j % 32;
// Eb/N0 from counter (g, 0, 2) when a range is given; Gaussian pair of
// codeword bits (2 v2, 2 v2 + 1) by Box-Muller from counter (g, v2, 1).
__device__ __forceinline__ void info_block(unsigned long long seed, unsigned long long g, int blk, uint32_t w[4])
{
    const uint4 r = philox4x32_10(make_uint4((uint32_t)g, (uint32_t)(g >> 32), (uint32_t)blk, 0u), (uint32_t)seed,
                                  (uint32_t)(seed >> 32));
    w[0] = r.x;
    w[1] = r.y;
    w[2] = r.z;
    w[3] = r.w;
}

__device__ __forceinline__ float frame_ebn0(unsigned long long seed, unsigned long long g, float lo, float hi)
{
    if (!(hi > lo)) return lo;
    const uint4 r = philox4x32_10(make_uint4((uint32_t)g, (uint32_t)(g >> 32), 0u, 2u), (uint32_t)seed,
                                  (uint32_t)(seed >> 32));
    return lo + (hi - lo) * (float)u01(r.x);
}

// LLRs (reference convention) of codeword bits b0, b1 at pair v2
__device__ __forceinline__ float2 noisy_pair(unsigned long long seed, unsigned long long g, int v2, uint8_t b0,
                                             uint8_t b1, double s2, double sig)
{
    const uint4 r = philox4x32_10(make_uint4((uint32_t)g, (uint32_t)(g >> 32), (uint32_t)v2, 1u), (uint32_t)seed,
                                  (uint32_t)(seed >> 32));
    double n0, n1;
    box_muller(r.x, r.y, n0, n1);
    const double y0 = (b0 ? -1.0 : 1.0) + sig * n0;
    const double y1 = (b1 ? -1.0 : 1.0) + sig * n1;
    // LLR = 2y/sigma^2 for bit 0 -> +1, negated to the reference convention
    float2 o;
    o.x = (float)(-2.0 * y0 / s2);
    o.y = (float)(-2.0 * y1 / s2);
    return o;
}

__device__ __forceinline__ double frame_s2(const EncodeDesc& d, float eb)
{
    const float rate = (float)d.k / (float)d.n;
    return 1.0 / (2.0 * (double)rate * pow(10.0, (double)eb / 10.0));
}

// One frame g of any code, generated by one workgroup (blockDim >= z):
// c is n + z bytes of LDS.  Writes llr_out[0..n), bits_out[0..n) (nullable).
__device__ __forceinline__ void generate_frame(const EncodeDesc& d, unsigned long long seed, unsigned long long g, float eb_lo,
                               float eb_hi, uint8_t* c, float* llr_out, uint8_t* bits_out, float* ebn0_out)
{
    const int z = d.z, k = d.k, n = d.n;
    const int tid = threadIdx.x;
    const int nb = (k + 127) / 128;
    for (int blk = tid; blk < nb; blk += blockDim.x) {
        uint32_t w[4];
        info_block(seed, g, blk, w);
        for (int q = 0; q < 128; q++) {
            const int j = blk * 128 + q;
            if (j < k) c[j] = (w[q >> 5] >> (q & 31)) & 1u;
        }
    }
    __syncthreads();
    const bool active = tid < z;
    encode_frame(d, c, c + n, active ? tid : 0, active, [&](int i, uint8_t bit) { c[k + i] = bit; });
    __syncthreads();
    const float eb = frame_ebn0(seed, g, eb_lo, eb_hi);
    const double s2 = frame_s2(d, eb);
    const double sig = sqrt(s2);
    if (tid == 0 && ebn0_out) *ebn0_out = eb;
    for (int v2 = tid; v2 < n / 2; v2 += blockDim.x) {
        const uint8_t b0 = c[2 * v2], b1 = c[2 * v2 + 1];
        reinterpret_cast<float2*>(llr_out)[v2] = noisy_pair(seed, g, v2, b0, b1, s2, sig);
        if (bits_out) {
            bits_out[2 * v2] = b0;
            bits_out[2 * v2 + 1] = b1;
        }
    }
}

}  // namespace wl
