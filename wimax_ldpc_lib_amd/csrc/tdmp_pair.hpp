// tdmp_pair.hpp -- small-batch shape of the TDMP normalised min-sum decoder:
// one frame per workgroup, two lanes per row.
//
// A batch of a few hundred or a few thousand frames gives the row-per-lane
// kernel (tdmp.hpp) about one wave per SIMD, and its time is then the length
// of one wave's serial chain: every layer reads, reduces and updates all D
// slots of the lane's row, 10 iterations of NL layers plus the syndrome
// checks.  Here lanes 2r and 2r + 1 of the workgroup share row r of every
// layer: the even lane takes slots 0 .. H-1 (H = ceil(D / 2)), the odd lane
// slots H .. D-1 (with D odd, its last slot is empty).  Each lane reduces its
// slots to (min1, min2, sign parity), the pair swaps the three values with
// one DPP quad permutation and merges them (the two smallest of four: one
// min, one min, one med3), and each lane updates its own slots.  The chain
// per layer is about half as long; the layer loop, the early-stop protocol
// and the outputs are those of the reference's decoder
// (lib/ldpc_decoder.cc:122-262), as in tdmp.hpp.
//
// Results are bit-identical to tdmp.hpp's: the merged (min1, min2) are the
// two smallest |L| of the row (clamped at 10000), the parity is the XOR of
// the row's sign bits, and every slot gets the same FP64 message step.
//
// The slot addresses and messages of a lane's half rows live in VGPRs (about
// 40 + 40 per lane), the frame's posteriors in LDS.  Contiguous frames only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qc_code.hpp"
#include "tdmp.hpp"

namespace wl {

constexpr int kPairMaxThreads = 192;   // 2 lanes x 96 rows

// threads of the workgroup: 2z lanes, whole waves
template <int Z>
constexpr int pair_threads()
{
    return ((2 * Z + 63) / 64) * 64;
}

// value of the adjacent lane (lane ^ 1): DPP quad permutation [1, 0, 3, 2]
// (every lane's source is inside its quad, so no bound control applies)
__device__ __forceinline__ uint32_t pair_swap(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
}

template <int RATE, int Z>
__global__ __launch_bounds__(kPairMaxThreads) void tdmp_pair_kernel(DecodeArgs a)
{
    using CD = Code<RATE, Z>;
    constexpr int NL = CD::NL;
    constexpr int N = CD::N;
    constexpr int HM = (CD::C + 1) / 2;   // slots per lane, largest layer
    constexpr int T = pair_threads<Z>();
    constexpr int NW = T / 64;
    static_assert(T <= kPairMaxThreads, "workgroup larger than the launch bound");
    static_assert(kLdsPad >= 64, "LDS pad holds the per-wave flags and the failing-row count");

    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* wflag = reinterpret_cast<uint32_t*>(smem);        // [NW] per-wave "some row fails"
    uint32_t* rows_failing = reinterpret_cast<uint32_t*>(smem + 32);
    float* rx = reinterpret_cast<float*>(smem + kLdsPad);

    const unsigned long long g = blockIdx.x;   // this workgroup's frame
    if (g >= a.nframes) return;                // whole workgroup, before any barrier
    const int tid = threadIdx.x;
    const bool active = tid < 2 * Z;
    const int q = active ? tid : 0;
    const int r = q >> 1;
    const bool odd = (q & 1) != 0;

    // the frame's LLRs into LDS: all 16-byte loads of a thread in flight at once
    {
        constexpr int NV = N / 4;
        constexpr int U = (NV + T - 1) / T;
        const f32x4* src = reinterpret_cast<const f32x4*>(a.llr + g * N);
        f32x4* dst = reinterpret_cast<f32x4*>(rx);
        f32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (tid + u * T < NV) v[u] = __builtin_nontemporal_load(src + tid + u * T);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (tid + u * T < NV) dst[tid + u * T] = v[u];
    }
    if (tid == 0) *rows_failing = 0;

    // LDS byte address of each half-row slot: variable blk * z + (r + s) mod z
    // of the slot this lane takes (the empty slot of an odd lane reads the even
    // lane's last slot and is never written)
    const uint32_t base = lds_addr(smem) + (uint32_t)kLdsPad;
    uint32_t ad[NL][HM];
    float msg[NL][HM];
    static_for<0, NL>([&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        constexpr int D = CD::deg(L);
        constexpr int H = (D + 1) / 2;
        static_for<0, H>([&](auto Kc) {
            constexpr int k = decltype(Kc)::value;
            constexpr int J1 = H + k < D ? H + k : k;
            constexpr int b0 = CD::blk(L, k), s0 = CD::sh(L, k);
            constexpr int b1 = CD::blk(L, J1), s1 = CD::sh(L, J1);
            const int b = odd ? b1 : b0;
            const int s = odd ? s1 : s0;
            const int c = r + s >= Z ? r + s - Z : r + s;
            uint32_t o = base + (uint32_t)((b * Z + c) * 4);
            asm volatile("" : "+v"(o));   // one register per slot, used as is by its reads and writes
            ad[L][k] = o;
            msg[L][k] = 0.0f;
        });
    });
    __syncthreads();

    // One layer (lib/ldpc_decoder.cc:143-240) for this lane's half of row r.
    auto layer = [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        constexpr int D = CD::deg(L);
        constexpr int H = (D + 1) / 2;
        constexpr bool PAD = (D & 1) != 0;   // the odd lane's slot H - 1 is empty
        float x[H];
        float mn1 = 10000.0f, mn2 = 10000.0f;
        uint32_t sx = 0;
        static_for<0, H>([&](auto Kc) {
            constexpr int k = decltype(Kc)::value;
            const float v = *lds_ptr(ad[L][k]) - msg[L][k];
            x[k] = v;
            float av = fabsf(v);
            uint32_t vb = __float_as_uint(v);
            if constexpr (PAD && k == H - 1) {
                av = odd ? 10000.0f : av;
                vb = odd ? 0u : vb;
            }
            // running two smallest: med3 of (min1, min2, |L|) is the new min2
            mn2 = __builtin_amdgcn_fmed3f(mn1, mn2, av);
            mn1 = fminf(mn1, av);
            sx ^= vb;
        });
        // merge with the other half of the row: the two smallest of
        // {mn1, mn2, p1, p2} (mn1 <= mn2, p1 <= p2) are min(mn1, p1) and
        // med3(mn1, p1, min(mn2, p2)).  Both pairs are non-negative floats,
        // ordered as their bit patterns: unsigned compares, which take the
        // swapped operands as they come (no float canonicalisation)
        {
            const uint32_t u1 = __float_as_uint(mn1), u2 = __float_as_uint(mn2);
            const uint32_t q1 = pair_swap(u1), q2 = pair_swap(u2);
            const uint32_t m = u2 < q2 ? u2 : q2;
            const uint32_t lo = u1 < q1 ? u1 : q1, hi = u1 < q1 ? q1 : u1;
            mn1 = __uint_as_float(lo);
            mn2 = __uint_as_float(hi < m ? hi : m);
        }
        sx ^= pair_swap(sx);
        // message sign bit of slot n = signbit(L_n) ^ parity(signbits) ^ (D & 1)
        const uint32_t kw = (sx ^ ((uint32_t)(D & 1) << 31)) & 0x80000000u;
        // s*(min*0.815) in double (:220-223), as tdmp.hpp
        const double d1 = __dmul_rn((double)mn1, 0.815);
        const double d2 = __dmul_rn((double)mn2, 0.815);
        const uint32_t h1 = (uint32_t)__double2hiint(d1) | kw, h2 = (uint32_t)__double2hiint(d2) | kw;
        const int l1 = __double2loint(d1), l2 = __double2loint(d2);
        static_for<0, H>([&](auto Kc) {
            constexpr int k = decltype(Kc)::value;
            const float v = x[k];
            const bool e = fabsf(v) == mn1;
            const uint32_t hi = (e ? h2 : h1) ^ (__float_as_uint(v) & 0x80000000u);
            const double dsg = __hiloint2double((int)hi, e ? l2 : l1);
            const float m = __double2float_rn(dsg);
            const float p = __double2float_rn(__dadd_rn((double)v, dsg));
            if constexpr (PAD && k == H - 1) {
                if (!odd) {
                    msg[L][k] = m;
                    *lds_ptr(ad[L][k]) = p;
                }
            } else {
                msg[L][k] = m;
                *lds_ptr(ad[L][k]) = p;
            }
        });
    };

    // Failing parity checks of row r (lib/ldpc_decoder.cc:67-104 over
    // hd = rx >= 0), counted on the even lane: each slot's compare is
    // ballotted, the masks XORed per layer, and row r's parity is bit 2r ^
    // bit 2r + 1 of the result.
    auto fail_rows = [&]() -> int {
        // all of the lane's slot reads in flight at once, then the compares
        float hv[NL][HM];
        static_for<0, NL>([&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            static_for<0, (CD::deg(L) + 1) / 2>([&](auto Kc) {
                hv[L][decltype(Kc)::value] = *lds_ptr(ad[L][decltype(Kc)::value]);
            });
        });
        static_for<0, NL>([&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            static_for<0, (CD::deg(L) + 1) / 2>([&](auto Kc) {
                float& h = hv[L][decltype(Kc)::value];
                asm volatile("" : "+v"(h));
            });
        });
        int fails = 0;
        static_for<0, NL>([&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            constexpr int D = CD::deg(L);
            constexpr int H = (D + 1) / 2;
            constexpr bool PAD = (D & 1) != 0;
            uint64_t pm = 0;
            static_for<0, H>([&](auto Kc) {
                constexpr int k = decltype(Kc)::value;
                bool b = hv[L][k] >= 0.0f;
                if constexpr (PAD && k == H - 1) b = b && !odd;
                pm ^= __builtin_amdgcn_ballot_w64(b);
            });
            const uint64_t rowp = (pm ^ (pm >> 1)) & 0x5555555555555555ull;
            fails += __builtin_amdgcn_inverse_ballot_w64(rowp) ? 1 : 0;
        });
        return fails;
    };
    // Whether some row of this wave's rows of layer L fails: row r's parity
    // is bit 2r ^ bit 2r + 1 of the XORed ballots, masked to the active rows
    // (a scalar test, no per-lane count).
    const uint64_t act_rows = __builtin_amdgcn_ballot_w64(active) & 0x5555555555555555ull;
    auto layer_fails = [&](auto Lc) -> bool {
        constexpr int L = decltype(Lc)::value;
        constexpr int D = CD::deg(L);
        constexpr int H = (D + 1) / 2;
        constexpr bool PAD = (D & 1) != 0;
        float hv[H];
        static_for<0, H>([&](auto Kc) { hv[decltype(Kc)::value] = *lds_ptr(ad[L][decltype(Kc)::value]); });
        uint64_t pm = 0;
        static_for<0, H>([&](auto Kc) {
            constexpr int k = decltype(Kc)::value;
            bool b = hv[k] >= 0.0f;
            if constexpr (PAD && k == H - 1) b = b && !odd;
            pm ^= __builtin_amdgcn_ballot_w64(b);
        });
        return ((pm ^ (pm >> 1)) & act_rows) != 0;
    };
    // does any row of the frame fail (uniform over the workgroup).  Only
    // this is needed for the stop decisions (compute_syndrome(hd, true)
    // returns at the first failing row), so layer 0 is checked first: a
    // frame with errors almost always fails there, and the other layers'
    // reads (all in flight at once) are issued only when it does not.
    auto frame_fails = [&]() -> bool {
        bool w = layer_fails(std::integral_constant<int, 0>{});
        if (!w) w = __builtin_amdgcn_ballot_w64(active && fail_rows() != 0) != 0;
        if constexpr (NW == 1) {
            return w;
        } else {
            if ((tid & 63) == 0) wflag[tid >> 6] = w ? 1u : 0u;
            __syncthreads();
            bool any = false;
#pragma unroll
            for (int i = 0; i < NW; i++) any = any || wflag[i] != 0;
            __syncthreads();   // the flags are read before the next check writes them
            return any;
        }
    };

    // Between layers: with several waves a workgroup barrier; with one
    // wave, program order alone (the LDS executes one wave's reads and
    // writes in issue order), so the next layer's reads follow this layer's
    // writes without waiting for them
    auto layer_sync = [&]() {
        if constexpr (NW == 1) {
            __builtin_amdgcn_wave_barrier();
            asm volatile("" ::: "memory");
        } else {
            __syncthreads();
        }
    };

    bool done = false;   // uniform: one frame per workgroup
    int it_done = 0;
    if (a.early_stop) done = !frame_fails();   // initial check (:130-139)
    for (int it = 0; it < a.max_iter && !done; it++) {
        static_for<0, NL>([&](auto Lc) {
            if (active) layer(Lc);
            layer_sync();
        });
        if (a.early_stop && !frame_fails()) {   // :242-251
            done = true;
            it_done = it + 1;
        }
    }
    const bool converged = a.early_stop && done;
    if (!converged) it_done = a.max_iter;
    // Final syndrome: the weight where it is asked for (a frame that stopped
    // early satisfies every check: weight 0); with fixed iterations the
    // return value only needs to know whether a row fails; with early stop
    // it follows from the stop decisions.
    if (a.synd) {
        const int fr = active && !converged ? fail_rows() : 0;
        if (fr) atomicAdd(rows_failing, (uint32_t)fr);
    } else if (a.ret && !a.early_stop) {
        if (frame_fails() && tid == 0) *rows_failing = 1;
    }
    __syncthreads();
    if (tid == 0) {
        const uint32_t w = *rows_failing;
        if (a.synd) a.synd[g] = (uint16_t)w;
        if (a.iters) a.iters[g] = (uint8_t)it_done;
        if (a.ret) {
            uint8_t rv;
            if (a.early_stop) rv = (converged || a.max_iter == 0) ? 0 : 1;   // :124,141,261
            else rv = w != 0;
            a.ret[g] = rv;
        }
    }
    if (a.post) {
        constexpr int NV = N / 4;
        const f32x4* src = reinterpret_cast<const f32x4*>(rx);
        f32x4* dst = reinterpret_cast<f32x4*>(a.post + g * N);
        for (int i = tid; i < NV; i += T) __builtin_nontemporal_store(src[i], dst + i);
    }
    if (a.hard) {
        // ballots of 64 consecutive variables, bytes written by lanes 0..7
        constexpr int kChunks = (N + 63) / 64;
        const int lane = tid & 63;
        for (int c = tid >> 6; c < kChunks; c += NW) {
            const int v = c * 64 + lane;
            const bool bit = v < N && rx[v] >= 0.0f;
            const unsigned long long m = __ballot(bit);
            const int byte = c * 8 + lane;
            if (lane < 8 && byte < N / 8) a.hard[g * (N / 8) + byte] = (uint8_t)(m >> (8 * lane));
        }
    }
}

template <int RATE, int Z>
hipError_t launch_tdmp_pair(const DecodeArgs& a, hipStream_t stream)
{
    if (a.nframes == 0) return hipSuccess;
    if (a.nframes > 0x7fffffffull) return hipErrorInvalidValue;
    if (a.frame_idx || a.bucket || a.frame_off || a.tile_counter) return hipErrorInvalidValue;   // contiguous only
    constexpr size_t lds = kLdsPad + (size_t)Code<RATE, Z>::N * sizeof(float);
    hipLaunchKernelGGL((tdmp_pair_kernel<RATE, Z>), dim3((unsigned)a.nframes), dim3(pair_threads<Z>()), lds, stream,
                       a);
    return hipGetLastError();
}

using PairLaunch = hipError_t (*)(const DecodeArgs&, hipStream_t);

// defined in decode_p<rate>.hip: the 19 launchers of one rate, by z index
#define WL_DECLARE_PAIR(R) const PairLaunch* pair_table_r##R();
WL_DECLARE_PAIR(0) WL_DECLARE_PAIR(1) WL_DECLARE_PAIR(2) WL_DECLARE_PAIR(3) WL_DECLARE_PAIR(4) WL_DECLARE_PAIR(5)
#undef WL_DECLARE_PAIR

inline PairLaunch pair_launcher(int rate, int z)
{
    if (!valid_z(z)) return nullptr;
    const PairLaunch* t = nullptr;
    switch (rate) {
    case 0: t = pair_table_r0(); break;
    case 1: t = pair_table_r1(); break;
    case 2: t = pair_table_r2(); break;
    case 3: t = pair_table_r3(); break;
    case 4: t = pair_table_r4(); break;
    case 5: t = pair_table_r5(); break;
    default: return nullptr;
    }
    return t[index_of_z(z)];
}

// frames up to which a decoder of lifting size z uses this kernel by
// default, measured on MI355X against the row-per-lane kernel: with early
// stop up to 32 waves per SIMD of num_cu compute units (4 SIMDs each) --
// each workgroup retires as soon as its one frame converges, which keeps
// this kernel ahead (rate 1/2 z=24: 347 vs 378 us at 65536 frames; rate 5/6
// z=96 even at 65536) -- and with a fixed iteration count up to 3 (rate 1/2
// z=24: 52 vs 54 us at 4096 frames, 92 vs 86 at 8192; rate 5/6 z=96: 47 vs
// 51 us at 1024, 80 vs 76 at 2048)
inline size_t pair_default_max_frames(int z, int num_cu, bool early_stop)
{
    const int waves = (2 * z + 63) / 64;
    return (size_t)num_cu * 4 * (early_stop ? 32 : 3) / (size_t)waves;
}

}  // namespace wl

#define WL_DEFINE_PAIR_TABLE(R)                                                                            \
    namespace wl {                                                                                        \
    const PairLaunch* pair_table_r##R()                                                                   \
    {                                                                                                     \
        static const PairLaunch t[kNumZ] = {                                                              \
            &launch_tdmp_pair<R, 24>, &launch_tdmp_pair<R, 28>, &launch_tdmp_pair<R, 32>,                 \
            &launch_tdmp_pair<R, 36>, &launch_tdmp_pair<R, 40>, &launch_tdmp_pair<R, 44>,                 \
            &launch_tdmp_pair<R, 48>, &launch_tdmp_pair<R, 52>, &launch_tdmp_pair<R, 56>,                 \
            &launch_tdmp_pair<R, 60>, &launch_tdmp_pair<R, 64>, &launch_tdmp_pair<R, 68>,                 \
            &launch_tdmp_pair<R, 72>, &launch_tdmp_pair<R, 76>, &launch_tdmp_pair<R, 80>,                 \
            &launch_tdmp_pair<R, 84>, &launch_tdmp_pair<R, 88>, &launch_tdmp_pair<R, 92>,                 \
            &launch_tdmp_pair<R, 96>};                                                                    \
        return t;                                                                                         \
    }                                                                                                     \
    }
