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

#include <type_traits>

#include "channel_dev.hpp"
#include "encoder.hpp"

namespace wl {

// One workgroup encodes a tile of F x passes frames: F frames at a time
// (F z <= 512 lanes), `passes` times.  The tile's info words and codewords
// are contiguous in HBM, so the systematic part is copied in vec-byte words
// (16, 4 or 1: the info bytes as given, lib/ldpc_encoder.cc:90) while the 0/1
// bits go to LDS; then every pass computes its frames' parity bits into LDS;
// then the parity parts are stored in vec-byte words.  A tile of several
// passes keeps a few loads per thread in flight.
// RATE < 0: the run-time descriptor's loops; otherwise code (RATE, Z) with
// its block columns and shifts as constants.
template <int V>
__device__ __forceinline__ void copy_info(const uint8_t* __restrict__ src, uint8_t* __restrict__ cw, uint8_t* u,
                                          int k, int n, int nbytes, unsigned long long frame0)
{
    using W = typename std::conditional<V == 16, uint4, typename std::conditional<V == 4, uint32_t, uint8_t>::type>::type;
    const int nw = nbytes / V;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) {
        const W w = reinterpret_cast<const W*>(src)[i];
        const int f = (V * i) / k, j = V * i - f * k;
        *reinterpret_cast<W*>(cw + (frame0 + f) * n + j) = w;
        // bytes != 0 -> 1: bit 7 of (b & 0x7f) + 0x7f, or of b itself
        auto nz = [](uint32_t x) { return ((((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) >> 7) & 0x01010101u; };
        if constexpr (V == 16) {
            reinterpret_cast<uint4*>(u)[i] = make_uint4(nz(w.x), nz(w.y), nz(w.z), nz(w.w));
        } else if constexpr (V == 4) {
            reinterpret_cast<uint32_t*>(u)[i] = nz(w);
        } else {
            u[i] = w != 0;
        }
    }
}

template <int V>
__device__ __forceinline__ void store_parity(const uint8_t* par, uint8_t* __restrict__ cw, int k, int n, int m,
                                             int Feff, unsigned long long frame0)
{
    using W = typename std::conditional<V == 16, uint4, typename std::conditional<V == 4, uint32_t, uint8_t>::type>::type;
    const int mw = m / V;
    for (int i = threadIdx.x; i < Feff * mw; i += blockDim.x) {
        const int f = i / mw, q = i - f * mw;
        reinterpret_cast<W*>(cw + (frame0 + f) * n + k)[q] = reinterpret_cast<const W*>(par)[f * mw + q];
    }
}

template <int RATE, int Z>
__global__ __launch_bounds__(512) void encode_kernel(EncodeDesc d, int F, int passes, const uint8_t* __restrict__ info,
                                                     uint8_t* __restrict__ cw, unsigned long long nframes, int vec)
{
    extern __shared__ __align__(16) uint8_t esm[];
    const int z = d.z, k = d.k, n = d.n, m = d.m;
    const int T = F * passes;
    const unsigned long long frame0 = (unsigned long long)blockIdx.x * T;
    const unsigned long long left = nframes - frame0;
    const int Feff = left < (unsigned long long)T ? (int)left : T;
    const int tid = threadIdx.x;
    uint8_t* u = esm;                         // [T][k] info bits
    uint8_t* par = esm + (size_t)T * k;       // [T][m] parity bits
    uint8_t* p0 = par + (size_t)T * m;        // [T][z] P_0 scratch
    const uint8_t* src = info + frame0 * k;
    if (vec == 16) copy_info<16>(src, cw, u, k, n, Feff * k, frame0);
    else if (vec == 4) copy_info<4>(src, cw, u, k, n, Feff * k, frame0);
    else copy_info<1>(src, cw, u, k, n, Feff * k, frame0);
    __syncthreads();
    const int t = tid % z;
    for (int pass = 0; pass < passes; pass++) {
        const int fs = pass * F + tid / z;
        const bool active = tid < F * z && fs < Feff;
        const int fsa = active ? fs : 0;
        uint8_t Vq[kMaxLayers];
        if constexpr (RATE < 0) encode_lambda(d, u + (size_t)fsa * k, t, active, Vq);
        else encode_lambda_ct<RATE, Z>(u + (size_t)fsa * k, t, active, Vq);
        encode_parity(d, Vq, p0 + (size_t)fsa * z, t, active,
                      [&](int i, uint8_t bit) { par[(size_t)fs * m + i] = bit; });
    }
    __syncthreads();
    if (vec == 16) store_parity<16>(par, cw, k, n, m, Feff, frame0);
    else if (vec == 4) store_parity<4>(par, cw, k, n, m, Feff, frame0);
    else store_parity<1>(par, cw, k, n, m, Feff, frame0);
}

// Bit-sliced encoder of the compiled codes.  A workgroup encodes a tile of
// 64 frames at once: bit f of every 64-bit word belongs to frame f of the
// tile, so the XOR steps of the encoder (lambda blocks, prediction and
// correction, back-substitution) run once for all 64 frames.
//   1. The tile's info bytes come in chunks of CB bytes per frame: coalesced
//      VEC-byte loads, copied to the codeword's systematic part (the info
//      bytes as given, lib/ldpc_encoder.cc:90) and staged in LDS.  Lane f of
//      a wave reads byte j of frame f, and one ballot of (byte != 0) is the
//      word of position j (the reference's `1 && infoword[...]`, :43),
//      stored to LDS by lane 0.
//   2. Lane t < z computes lambda_q[t] of every base row from the words
//      (:29-48) and the parity words (:55-103), as encode_parity does for one
//      frame.
//   3. Lane (fs, q) of a wave holds the parity words of VEC positions and
//      writes, for frames 4g + fs, the 0/1 bytes of its frame's bits as one
//      VEC-byte store: 16 lanes store 16 VEC consecutive bytes of one frame,
//      whole cache lines (4-byte stores, two half lines per line, measured
//      12 % more HBM write traffic).
// This trades the per-frame, per-position work of encode_kernel (one byte
// per info slot, position and frame) for two byte<->bit transposes of about
// 2 and 1.75 VALU operations per 64 frame-bits.
constexpr int kBsFrames = 64;
constexpr int kBsThreads = 256;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// bytes per frame staged per step: the largest multiple of VEC that divides
// k and is at most 128 (four waves of at most 32 positions each)
template <int K, int VEC>
constexpr int bs_chunk()
{
    int c = 128 - 128 % VEC;
    while (c > VEC && K % c != 0) c -= VEC;
    return c;
}

// staging row stride: a multiple of 4 bytes, an odd number of dwords, so the
// 64 lanes of a byte read (one row each) hit 64 different banks
template <int CB>
constexpr int bs_row()
{
    return CB + 4;
}

// access width of the compiled code (rate, z): 16 bytes where k % 16 == 0,
// else 8 (k = kb z with z % 4 == 0 is always a multiple of 8)
template <int RATE, int Z>
constexpr int bs_vec()
{
    return Code<RATE, Z>::K % 16 == 0 ? 16 : 8;
}

// LDS of a tile: info words [K] and P_0 [Z], then one region holding the
// staged bytes (phase 1) and later the parity words [M] (phases 2 and 3)
template <int RATE, int Z>
constexpr size_t bs_lds()
{
    using CD = Code<RATE, Z>;
    constexpr int CB = bs_chunk<CD::K, bs_vec<RATE, Z>()>();
    constexpr size_t stg = (size_t)kBsFrames * bs_row<CB>(), par = (size_t)CD::M * sizeof(uint64_t);
    return (size_t)(CD::K + Z) * sizeof(uint64_t) + (stg > par ? stg : par);
}

template <int RATE, int Z>
__global__ __launch_bounds__(kBsThreads) void encode_bs_kernel(EncodeDesc d, const uint8_t* __restrict__ info,
                                                               uint8_t* __restrict__ cw, unsigned long long nframes)
{
    using CD = Code<RATE, Z>;
    constexpr int K = CD::K, M = CD::M, N = CD::N, KB = 24 - CD::NL;
    constexpr int VEC = bs_vec<RATE, Z>();
    constexpr int CB = bs_chunk<K, VEC>();
    constexpr int SROW = bs_row<CB>();
    constexpr int WPF = CB / VEC;              // loads per frame and chunk
    constexpr int PPW = CB / 4;                // positions per wave and chunk
    static_assert(K % CB == 0 && CB % VEC == 0 && CB % 4 == 0 && PPW <= 32, "chunk shape");
    static_assert(M % 4 == 0 && N % 16 == 0, "parity quads and codeword rows");
    using Word = typename std::conditional<VEC == 16, u32x4, u32x2>::type;

    extern __shared__ __align__(16) uint8_t esm[];
    uint64_t* wl = reinterpret_cast<uint64_t*>(esm);    // [K] info words
    uint64_t* p0 = wl + K;                               // [Z] P_0 scratch
    uint8_t* stg = reinterpret_cast<uint8_t*>(p0 + Z);   // [64][SROW] staged bytes (phase 1)
    uint64_t* pl = reinterpret_cast<uint64_t*>(stg);     // [M] parity words (phases 2, 3)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long frame0 = (unsigned long long)blockIdx.x * kBsFrames;
    const unsigned long long left = nframes - frame0;
    const int fe = left < (unsigned long long)kBsFrames ? (int)left : kBsFrames;

    // 1. info bytes -> systematic part + info words.  The loads of the next
    // chunk are issued before the current chunk is transposed, so they are in
    // flight across its barriers.  (Rows of frames >= fe are left as they
    // are: their bits never leave the tile.)
    constexpr int LPT = (kBsFrames * WPF + kBsThreads - 1) / kBsThreads;   // loads per thread and chunk
    Word v[LPT];
    auto load = [&](int j0) {
#pragma unroll
        for (int u = 0; u < LPT; u++) {
            const int i = tid + u * kBsThreads, f = i / WPF, w = i - f * WPF;
            if (i < kBsFrames * WPF && f < fe)
                v[u] = __builtin_nontemporal_load(reinterpret_cast<const Word*>(info + (frame0 + f) * K + j0) + w);
        }
    };
    load(0);
    for (int j0 = 0; j0 < K; j0 += CB) {
#pragma unroll
        for (int u = 0; u < LPT; u++) {
            const int i = tid + u * kBsThreads, f = i / WPF, w = i - f * WPF;
            if (i < kBsFrames * WPF && f < fe) {
                __builtin_nontemporal_store(v[u], reinterpret_cast<Word*>(cw + (frame0 + f) * N + j0) + w);
                uint32_t* s = reinterpret_cast<uint32_t*>(stg + f * SROW + VEC * w);
                s[0] = v[u].x;
                s[1] = v[u].y;
                if constexpr (VEC == 16) {
                    s[2] = v[u].z;
                    s[3] = v[u].w;
                }
            }
        }
        __syncthreads();
        if (j0 + CB < K) load(j0 + CB);
        // all of the lane's byte reads first, then the ballots in groups of
        // eight, written by lane 0 (one exec switch per group)
        uint32_t b[PPW];
#pragma unroll
        for (int p = 0; p < PPW; p++) b[p] = stg[lane * SROW + wave * PPW + p];
#pragma unroll
        for (int g0 = 0; g0 < PPW; g0 += 8) {
            uint64_t m[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (g0 + u < PPW) m[u] = __builtin_amdgcn_ballot_w64(b[g0 + u] != 0);
            if (lane == 0) {
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (g0 + u < PPW) wl[j0 + wave * PPW + g0 + u] = m[u];
            }
        }
        __syncthreads();
    }

    // 2. lambda and parity words, lane t < z
    const bool act = tid < Z;
    const int t = act ? tid : 0;
    uint64_t V[kMaxLayers];
#pragma unroll
    for (int q = 0; q < kMaxLayers; q++) V[q] = 0;
    if (act) {
        static_for<0, CD::NL>([&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            uint64_t v = 0;
            static_for<0, CD::deg(L)>([&](auto Jc) {
                constexpr int J = decltype(Jc)::value;
                constexpr int b = CD::blk(L, J), s = CD::sh(L, J);
                if constexpr (b < KB) {
                    const int p = t >= Z - s ? t + s - Z : t + s;
                    v ^= wl[b * Z + p];
                }
            });
            V[L] = v;
        });
    }
    encode_parity(d, V, p0, t, act, [&](int i, uint64_t w) { pl[i] = w; });
    __syncthreads();

    // 3. parity words -> 0/1 bytes of the parity part.  Lane (fs, q) holds
    // the words of positions i0 .. i0 + VEC - 1 of a block of 16 VEC
    // positions and writes them for frames 4 g + fs as one VEC-byte store, so
    // 16 lanes write 16 VEC consecutive bytes of a frame (whole cache lines).
    const int fs = lane >> 4, q = lane & 15;
    constexpr int PB = 16 * VEC;
    static_assert(M % VEC == 0, "parity part in whole lanes");
    for (int i0 = VEC * q; i0 < (M + PB - 1) / PB * PB; i0 += PB) {
        const bool qa = i0 < M;
        uint32_t lo[VEC], hi[VEC];
#pragma unroll
        for (int e = 0; e < VEC; e += 2) {
            u32x4 x = {0, 0, 0, 0};
            if (qa) x = *reinterpret_cast<const u32x4*>(pl + i0 + e);
            lo[e] = x.x;
            hi[e] = x.y;
            lo[e + 1] = x.z;
            hi[e + 1] = x.w;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int f = 4 * (wave + 4 * u) + fs;   // frames 0..31 for u < 2, 32..63 after
            const uint32_t* w = u < 2 ? lo : hi;
            const uint32_t sh = (uint32_t)(f & 31);
            uint32_t o[VEC / 4];
#pragma unroll
            for (int c = 0; c < VEC / 4; c++) {
                uint32_t x = __builtin_amdgcn_ubfe(w[4 * c], sh, 1);
                x |= __builtin_amdgcn_ubfe(w[4 * c + 1], sh, 1) << 8;
                x |= __builtin_amdgcn_ubfe(w[4 * c + 2], sh, 1) << 16;
                x |= __builtin_amdgcn_ubfe(w[4 * c + 3], sh, 1) << 24;
                o[c] = x;
            }
            if (qa && f < fe) {
                Word out;
                if constexpr (VEC == 16) out = Word{o[0], o[1], o[2], o[3]};
                else out = Word{o[0], o[1]};
                __builtin_nontemporal_store(out, reinterpret_cast<Word*>(cw + (frame0 + f) * N + K + i0));
            }
        }
    }
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
    d.compiled = 1;
    d.rate = rate;
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

using EncodeKernel = void (*)(EncodeDesc, int, int, const uint8_t*, uint8_t*, unsigned long long, int);

template <int... I>
constexpr auto make_encode_kernels(std::integer_sequence<int, I...>)
{
    struct Arr { EncodeKernel k[sizeof...(I)]; };
    return Arr{{&encode_kernel<I / kNumZ, z_of_index(I % kNumZ)>...}};
}
// [rate * 19 + z index]
constexpr auto kEncodeKernelArr = make_encode_kernels(std::make_integer_sequence<int, kNumRates * kNumZ>{});
constexpr const EncodeKernel* kEncodeKernels = kEncodeKernelArr.k;

using EncodeBsKernel = void (*)(EncodeDesc, const uint8_t*, uint8_t*, unsigned long long);
struct EncodeBsEntry {
    EncodeBsKernel kern;
    int vec;      // alignment the info and codeword pointers need
    size_t lds;
};

template <int... I>
constexpr auto make_encode_bs_kernels(std::integer_sequence<int, I...>)
{
    struct Arr { EncodeBsEntry e[sizeof...(I)]; };
    return Arr{{EncodeBsEntry{&encode_bs_kernel<I / kNumZ, z_of_index(I % kNumZ)>,
                              bs_vec<I / kNumZ, z_of_index(I % kNumZ)>(),
                              bs_lds<I / kNumZ, z_of_index(I % kNumZ)>()}...}};
}
// [rate * 19 + z index]
constexpr auto kEncodeBsArr = make_encode_bs_kernels(std::make_integer_sequence<int, kNumRates * kNumZ>{});

hipError_t launch_encode(const EncodeDesc& d, const uint8_t* info, uint8_t* cw,
                         unsigned long long nframes, hipStream_t s)
{
    if (nframes == 0) return hipSuccess;
    const uintptr_t al = (uintptr_t)info | (uintptr_t)cw;
    // compiled codes: the bit-sliced kernel, 64 frames per workgroup, from
    // one full tile on (a smaller batch leaves most bits of every word
    // unused, and its chunk steps, one load round trip and two barriers each,
    // make it slower for a single frame than the byte-per-frame kernel, whose
    // loads are all in flight at once)
    if (d.compiled && d.rate >= 0 && d.rate < kNumRates && valid_z(d.z) && nframes >= (unsigned long long)kBsFrames) {
        const EncodeBsEntry& e = kEncodeBsArr.e[d.rate * kNumZ + index_of_z(d.z)];
        if (al % (uintptr_t)e.vec == 0) {
            const unsigned long long grid = (nframes + kBsFrames - 1) / kBsFrames;
            if (grid > 0x7fffffffull) return hipErrorInvalidValue;
            hipLaunchKernelGGL(e.kern, dim3((unsigned)grid), dim3(kBsThreads), e.lds, s, d, info, cw, nframes);
            return hipGetLastError();
        }
    }
    // F frames per pass (F z <= 512 lanes), passes per tile while the tile's
    // LDS (n + z bytes per frame) stays within 48 KiB
    int F = 512 / d.z;
    if (F < 1) F = 1;
    const size_t per = (size_t)d.n + d.z;
    while (F > 1 && (size_t)F * per > 49152) F--;
    int passes = 1;
    while (passes < 8 && (size_t)F * (passes + 1) * per <= 49152) passes++;
    const unsigned long long T = (unsigned long long)F * passes;
    const unsigned long long grid = (nframes + T - 1) / T;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    const size_t lds = (size_t)T * per;
    int vec = 1;
    if (d.k % 16 == 0 && d.n % 16 == 0 && d.m % 16 == 0 && al % 16 == 0) vec = 16;
    else if (d.k % 4 == 0 && d.n % 4 == 0 && d.m % 4 == 0 && al % 4 == 0) vec = 4;
    EncodeKernel kern = &encode_kernel<-1, 0>;
    if (d.compiled && d.rate >= 0 && d.rate < kNumRates && valid_z(d.z))
        kern = kEncodeKernels[d.rate * kNumZ + index_of_z(d.z)];
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(((F * d.z + 63) / 64) * 64), lds, s, d, F, passes, info, cw,
                       nframes, vec);
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
