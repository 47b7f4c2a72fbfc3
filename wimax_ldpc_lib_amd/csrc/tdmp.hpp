// tdmp.hpp -- layered (TDMP) min-sum QC-LDPC decoder kernel for gfx950.
//
// Restates the decoding loop of the reference (lib/ldpc_decoder.cc:122-262)
// on the GPU, bit-exact:
//   * one lane owns expanded row r of every layer of one frame; the z rows of a
//     layer are independent (disjoint variables), so a layer is one parallel
//     step and layers follow in base-row order, as the row loop at :143 does;
//   * the frame's posterior LLRs stay in LDS for all iterations; slot j of row
//     r reads variable blk*z + (r + s) mod z, i.e. rotated LDS indexing with
//     the shift folded into the instruction offset;
//   * the check-to-variable messages of a lane's rows live in VGPRs
//     (sum of layer degrees per lane, 76..88 floats);
//   * min1 / min2 / sign parity are reduced in registers per lane: min1 / min2
//     are the two smallest |L| of the row (min3 / med3 over slot triples,
//     merged into the running pair), and an edge receives min2 iff |L| == min1, which
//     equals the reference's first-strict-argmin choice (ties give min2 == min1);
//   * the normalised message step runs in FP64 exactly as the reference's
//     float*double expression (:220-223): one rounding to float of
//     s*(min*0.815) for the message and of (double)L + s*(min*0.815) for the
//     posterior;
//   * HBM is touched only by the coalesced LLR load, the packed hard-decision
//     store (ballot of rx >= 0, the VOLK slicer contract at :130,242) and the
//     optional posterior / per-frame status stores.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qc_code.hpp"

namespace wl {

struct DecodeArgs {
    const float* llr;          // [nframes][n] input LLRs, reference convention (>= 0 -> bit 1)
    float* post;               // [nframes][n] posteriors out (nullable)
    uint8_t* hard;             // [nframes][n/8] hard decisions, bit v%8 of byte v/8 (nullable)
    uint8_t* ret;              // [nframes] ldpc_decoder::decode return value (nullable)
    uint8_t* iters;            // [nframes] iterations run (nullable)
    uint16_t* synd;            // [nframes] failing parity checks of the output (nullable)
    const float* offsets;      // learned mode: [T][NL][C] per-block offsets
    const uint32_t* frame_idx; // nullable: slot i decodes frame frame_idx[bucket[0] + i]
    const uint32_t* bucket;    // nullable: {first slot, slot count}, read on the device
    const unsigned long long* frame_off;  // nullable: element offset of frame g in llr/post
                                          // (hard: frame_off[g] / 8 bytes); default g * n
    unsigned long long* tile_counter;     // nullable: persistent mode, tiles taken from this counter
    int persistent_grid;                  // persistent mode: workgroups launched (<= tiles)
    int fallback_iters;                   // MODE_LEARNED_FB: normalised min-sum iterations of re-queued frames
    uint8_t* fallback;                    // MODE_LEARNED_FB: [nframes] 1 = re-decoded (nullable)
    unsigned long long nframes;
    int frames_per_wg;         // F: frames decoded together by one workgroup
    int max_iter;
    int early_stop;            // 1: syndrome check before/after every iteration (as shipped)
    int n_offset_iters;        // T (learned mode)
    int rule;                  // MODE_ALT: 0 offset min-sum, 1 plain min-sum
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

// LDS byte address of a shared-memory pointer, and back: the address arithmetic
// of the rotated row reads then stays in 32-bit LDS space, where a constant
// term folds into the ds instruction's offset field.  The accesses are
// volatile, which keeps them in program order (as the layers issue them
// anyway) and stops the compiler from pairing two slots into one
// ds_read2 / ds_write2: the pair's 8-bit offset fields reach 1020 bytes, so
// pairing slots further apart cost a vector add for a new base address, and
// the kernels are bound by vector-ALU issue, not LDS issue (rate 5/6 z=96:
// 76 adds fewer per iteration, 1.8 % faster).
typedef volatile __attribute__((address_space(3))) float lds_float;
__device__ __forceinline__ uint32_t lds_addr(unsigned char* p)
{
    return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) unsigned char*)p;
}
__device__ __forceinline__ lds_float* lds_ptr(uint32_t a)
{
    return (lds_float*)(uintptr_t)a;
}

constexpr int gcd_int(int a, int b) { return b == 0 ? a : gcd_int(b, a % b); }

// Lane masks of the wrapped row reads: wave w holds threads 64w .. 64w+63,
// thread t owns row t mod Z, and slot (l, j) of that row wraps when
// row >= Z - shift(l, j).  The pattern repeats every P = Z / gcd(64, Z) waves.
template <int RATE, int Z>
struct WrapMasks {
    using CD = Code<RATE, Z>;
    static constexpr int P = Z / gcd_int(64, Z);
    struct Table {
        uint64_t m[P][CD::NL][CD::C];
    };
    static constexpr Table make()
    {
        Table t{};
        for (int p = 0; p < P; p++)
            for (int l = 0; l < CD::NL; l++)
                for (int j = 0; j < CD::C; j++) {
                    uint64_t m = 0;
                    if (j < CD::deg(l) && CD::sh(l, j) != 0)
                        for (int lane = 0; lane < 64; lane++)
                            if ((64 * p + lane) % Z >= Z - CD::sh(l, j)) m |= 1ull << lane;
                    t.m[p][l][j] = m;
                }
        return t;
    }
    static constexpr Table table = make();
};

// the mask table seen through a constant-address-space pointer: a pointer
// made opaque with an empty asm keeps the masks as scalar loads at their
// point of use instead of constants hoisted out of the loops
typedef const __attribute__((address_space(4))) uint64_t cmask_t;

// this lane's bit of a wave-uniform mask set ? b : a (one v_cndmask on the
// mask's scalar registers)
__device__ __forceinline__ uint32_t select_by_mask(uint32_t a, uint32_t b, uint64_t mask)
{
    return __builtin_amdgcn_inverse_ballot_w64(mask) ? b : a;
}

constexpr int kLdsPad = 768;   // per-frame counters and offsets; also keeps wrapped row bases >= 0
constexpr int kFrameOffOff = 512;   // extended tiles: element offset of each slot's frame (u64 per slot)
constexpr int kFrameIdOff = 680;    // extended tiles: global frame of each slot (u32 per slot)
constexpr int kTileSlot = 504; // persistent mode: the tile index broadcast, inside the pad
constexpr int kQueueOff = 128;      // MODE_LEARNED_FB: re-queue of failed frame ids (2F uint32)
constexpr int kFailOff = 304;       // MODE_LEARNED_FB: per-slot return values (F bytes)
constexpr int kQueueCountOff = 496; // MODE_LEARNED_FB: queue length
constexpr int kMaxThreads = 512;
constexpr int kMaxLdsBytes = 65536;

// MODE_LEARNED_FB: learned decoding, then frames whose syndrome fails are
// re-queued (per workgroup, in LDS) and decoded again from their input LLRs
// with normalised min-sum in the same persistent kernel.
// MODE_ALT: the reference's two other compile-time message rules
// (lib/ldpc_decoder.cc:20-21,225-238), chosen per launch by DecodeArgs::rule:
// offset min-sum s * max(min - 0.10, 0) in double, or plain min-sum s * min
// in float.
enum { MODE_NMS = 0, MODE_LEARNED = 1, MODE_LEARNED_FB = 2, MODE_ALT = 3 };

// Code-shape variants (bit mask), all bit-identical in results:
//   VAR_NOHOIST  recompute each slot's wrapped LDS address inside the layer
//                (2 VALU per slot) instead of keeping one address register per
//                distinct shift live across the iteration loop -> fewer VGPRs
//   VAR_REREAD   reload rx and recompute L = rx - msg in the second pass
//                instead of keeping the row's L values live -> fewer VGPRs
//   VAR_SEL      with VAR_NOHOIST: select between the two opaque row bases
//                (r and r - Z) per slot, shift folded into the LDS offset
//                (compare + select, 2 VALU per slot instead of 3)
//   VAR_SMASK    with VAR_SEL: the per-slot "row wraps" lane masks come from a
//                compile-time table (scalar loads) instead of a compare per
//                slot, so the select is the only VALU op of the address
//   VAR_W2       with VAR_SEL, two frames per workgroup and z % 32 == 0: lane
//                l of wave w owns row 32 w + l % 32 of frame l / 32, and each
//                wave runs its own compile-time copy of the iteration loop.
//                A slot's wrap boundary z - shift then falls inside at most
//                one wave's 32 rows, so the other waves address it without a
//                select.  Frames sit 32 floats apart beyond n in LDS so the
//                two halves of a wave read different banks.
//   VAR_W4       the same with four frames per workgroup and z % 16 == 0:
//                lane l of wave w owns row 16 w + l % 16 of frame l / 16
//                (frames 16 floats apart beyond n)
//   VAR_OCC4     no change to the code: the plain (contiguous) kernel is
//                compiled for 4 waves per SIMD (128 registers, spilling what
//                does not fit) instead of the registers the shape takes
enum { VAR_NOHOIST = 1, VAR_REREAD = 2, VAR_SEL = 4, VAR_SMASK = 8, VAR_W2 = 16, VAR_W4 = 32, VAR_OCC4 = 64 };
// the low-register code shape of the tuned variants
constexpr int VAR_LEAN = VAR_NOHOIST | VAR_SEL | VAR_SMASK;
constexpr int VAR_LEAN_W2 = VAR_LEAN | VAR_W2;
constexpr int VAR_LEAN_W4 = VAR_LEAN | VAR_W4;
// the hoisted-address shape compiled for 4 waves per SIMD
constexpr int VAR_HOIST4 = VAR_OCC4;

// rows of one frame per wave in the wave-split shapes (0: not split); the
// workgroup then decodes 64 / rows frames
template <int VAR>
constexpr int split_rows()
{
    return (VAR & VAR_W2) != 0 ? 32 : ((VAR & VAR_W4) != 0 ? 16 : 0);
}

// floats between consecutive frames in LDS
template <int N, int VAR>
constexpr int lds_frame_stride()
{
    return N + split_rows<VAR>();
}

template <int RATE, int Z>
constexpr int max_frames_per_wg()
{
    int f = kMaxThreads / Z;
    while (f > 1 && kLdsPad + f * 24 * Z * 4 > kMaxLdsBytes) f--;
    return f;
}

// One tile: frames [frame0, frame0 + Feff) of the (possibly indirect) frame
// list, decoded by one workgroup.  All threads of the workgroup call it.
// ids (nullable): explicit frame ids of the tile's slots (overrides the frame
// list); fail_out (nullable): per-slot return value, for re-queueing.
// EXT = false compiles the plain form only (contiguous frames, none of the
// indirection below), which keeps scalar-register pressure of the headline
// kernel at its minimum.
template <int RATE, int Z, int MODE, int VAR, bool EXT>
__device__ __forceinline__ void tdmp_tile(const DecodeArgs& a, unsigned char* smem, unsigned long long frame0,
                                          int Feff, const uint32_t* ids = nullptr, uint8_t* fail_out = nullptr)
{
    using CD = Code<RATE, Z>;
    constexpr int NL = CD::NL;
    constexpr int N = CD::N;
    constexpr int C = CD::C;
    constexpr int RS = split_rows<VAR>();   // wave-split shapes: rows of one frame per wave
    constexpr bool SPLIT = RS != 0;
    static_assert(!SPLIT || (Z % RS == 0 && (VAR & VAR_SEL) != 0), "wave-split shapes need z % rows == 0 and VAR_SEL");
    constexpr int NS = lds_frame_stride<N, VAR>();

    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem);
    float* rxall = reinterpret_cast<float*>(smem + kLdsPad);

    const int tid = threadIdx.x;
    const int F = a.frames_per_wg;
    int f, r;
    if constexpr (SPLIT) {
        f = (tid & 63) / RS;
        r = (tid >> 6) * RS + (tid & (RS - 1));
    } else {
        f = tid / Z;
        r = tid - f * Z;
    }
    const bool active = f < Feff;
    // wave phase of the wrap-mask table (uniform over the wave)
    const int wphase = (__builtin_amdgcn_readfirstlane(tid) >> 6) % WrapMasks<RATE, Z>::P;
    // global frame of slot s
    auto gframe = [&](int s) -> unsigned long long {
        if constexpr (EXT) {
            if (ids) return (unsigned long long)ids[s];
            const unsigned long long slot0 = a.bucket ? (unsigned long long)a.bucket[0] : 0ull;
            return a.frame_idx ? (unsigned long long)a.frame_idx[slot0 + frame0 + s] : frame0 + s;
        } else {
            return frame0 + s;
        }
    };
    bool contiguous = true;
    if constexpr (EXT) contiguous = !a.frame_idx && !a.frame_off && !ids;

    if (tid < F) cnt[tid] = 0;
    // Extended tiles: each slot's global frame and element offset, read once
    // (the frame list and offset table are a chain of dependent global
    // loads) and kept in LDS for the frame loads and the output stores.
    unsigned long long* foff = reinterpret_cast<unsigned long long*>(smem + kFrameOffOff);
    uint32_t* fid = reinterpret_cast<uint32_t*>(smem + kFrameIdOff);
    if constexpr (EXT) {
        static_assert(kFrameOffOff + 8 * (kMaxThreads / 24) <= kFrameIdOff, "slot offsets overlap the frame ids");
        static_assert(kFrameIdOff + 4 * (kMaxThreads / 24) <= kLdsPad, "slot frame ids overflow the LDS pad");
        if (!contiguous && tid < Feff) {
            // a frame list or queue holds 32-bit frame numbers
            const unsigned long long g = gframe(tid);
            fid[tid] = (uint32_t)g;
            foff[tid] = a.frame_off ? a.frame_off[g] : g * N;
        }
    }
    auto fo = [&](int s) -> unsigned long long {
        if constexpr (EXT) {
            if (!contiguous) return foff[s];
        }
        return (frame0 + s) * N;
    };
    if (contiguous) {
        const f32x4* src = reinterpret_cast<const f32x4*>(a.llr + frame0 * N);
        f32x4* dst = reinterpret_cast<f32x4*>(rxall);
        const int total4 = Feff * (N / 4);
        const int bd = blockDim.x;
        // up to 8 loads per thread in flight before the LDS stores
        for (int i0 = tid; i0 < total4; i0 += 8 * bd) {
            f32x4 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (i0 + u * bd < total4) v[u] = __builtin_nontemporal_load(src + i0 + u * bd);
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = i0 + u * bd;
                if (i < total4) {
                    if constexpr (NS == N) {
                        dst[i] = v[u];
                    } else {
                        const int s = i / (N / 4);
                        dst[i + s * ((NS - N) / 4)] = v[u];
                    }
                }
            }
        }
    } else {
        __syncthreads();   // slot offsets in LDS
        constexpr int per = N / 4;
        const int total4 = Feff * per;
        const int bd = blockDim.x;
        f32x4* dst = reinterpret_cast<f32x4*>(rxall);
        // all frames of the tile at once, up to 8 loads per thread in flight
        for (int i0 = tid; i0 < total4; i0 += 8 * bd) {
            f32x4 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = i0 + u * bd;
                if (i < total4) {
                    const int s = i / per;
                    v[u] = reinterpret_cast<const f32x4*>(a.llr + foff[s])[i - s * per];
                }
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = i0 + u * bd;
                if (i < total4) {
                    const int s = i / per;
                    dst[i + s * ((NS - N) / 4)] = v[u];
                }
            }
        }
    }
    __syncthreads();

    // Row base pointers: slot (l, j) of this lane's row reads
    // rx[blk*Z + r + s] or, once r + s wraps, rx[blk*Z + r + s - Z].
    float* const frame_base = rxall + (active ? f : 0) * NS;
    float* const base0 = frame_base + r;
    float* const baseW = base0 - Z;

    auto at = [&](auto Lc, auto Jc) -> float* {
        constexpr int L = decltype(Lc)::value;
        constexpr int J = decltype(Jc)::value;
        constexpr int b = CD::blk(L, J);
        constexpr int s = CD::sh(L, J);
        if constexpr (s == 0) {
            return base0 + b * Z;
        } else {
            return ((r >= Z - s) ? baseW : base0) + (b * Z + s);
        }
    };
    auto at_r = [&](auto Lc, auto Jc, int rr) -> float* {
        constexpr int L = decltype(Lc)::value;
        constexpr int J = decltype(Jc)::value;
        constexpr int b = CD::blk(L, J);
        constexpr int s = CD::sh(L, J);
        float* const b0 = frame_base + rr;
        if constexpr (s == 0) {
            return b0 + b * Z;
        } else {
            return ((rr >= Z - s) ? b0 - Z : b0) + (b * Z + s);
        }
    };

    float msg[NL][C];
    static_for<0, NL>([&](auto Lc) {
        static_for<0, C>([&](auto Jc) { msg[decltype(Lc)::value][decltype(Jc)::value] = 0.0f; });
    });

    // LDS byte address of slot (L, J) of this lane's row: one of the two row
    // bases (ob0: rows r, obW: r - Z) plus the slot's constant offset.  With
    // a wave-split shape, W is this wave's index (rows RS W .. RS W + RS - 1):
    // where all of them wrap or none does, the base is chosen at compile
    // time; otherwise (and without a wave-split shape) one select on the
    // lane mask.
    auto sel_addr = [&](auto Lc2, auto Jc2, auto Wc, uint32_t ob0, uint32_t obW, cmask_t* mt) -> uint32_t {
        constexpr int L2 = decltype(Lc2)::value;
        constexpr int J2 = decltype(Jc2)::value;
        constexpr int b = CD::blk(L2, J2);
        constexpr int s = CD::sh(L2, J2);
        constexpr uint32_t off = (uint32_t)((b * Z + s) * 4);
        if constexpr (s == 0) {
            return ob0 + off;
        } else if constexpr (SPLIT) {
            constexpr int lo = RS * decltype(Wc)::value, bnd = Z - s;   // rows >= bnd wrap
            if constexpr (bnd >= lo + RS) {
                return ob0 + off;
            } else if constexpr (bnd <= lo) {
                return obW + off;
            } else {
                // lanes whose row (l % RS) + lo >= bnd, in each of the 64 / RS frames
                constexpr uint64_t seg = ((RS == 32 ? 0xffffffffull : 0xffffull) << (bnd - lo)) &
                                         (RS == 32 ? 0xffffffffull : 0xffffull);
                uint64_t m = 0;
                for (int q = 0; q < 64 / RS; q++) m |= seg << (q * RS);
                // in the extended kernels (more scalar state) materialised
                // here, two scalar moves, rather than hoisted and spilled
                if constexpr (EXT) asm volatile("" : "+s"(m));
                return select_by_mask(ob0, obW, m) + off;
            }
        } else {
            return select_by_mask(ob0, obW, mt[J2]) + off;
        }
    };

    // The row as a register value is only needed by the compare-based slot
    // addressing (shapes without a mask table); the mask-table and wave-split
    // shapes address slots from the two row bases alone.
    constexpr bool uses_rr = ((VAR & VAR_SEL) != 0) ? ((VAR & VAR_SMASK) == 0 && !SPLIT) : ((VAR & VAR_NOHOIST) != 0);
    // VAR_SEL: LDS byte addresses of the two row bases, made opaque again at
    // every layer; each layer's asm consumes the previous layer's value, so no
    // copy is kept live across the loop
    uint32_t lob0 = lds_addr(smem) + (uint32_t)(kLdsPad + ((active ? f : 0) * NS + r) * 4);
    uint32_t lobW = lob0 - (uint32_t)(Z * 4);

    // One layer for this lane's row (lib/ldpc_decoder.cc:143-240).
    auto layer = [&](auto Lc, int t, auto Wc) {
        constexpr int L = decltype(Lc)::value;
        constexpr int D = CD::deg(L);
        int rr = r;
        if constexpr (uses_rr) asm volatile("" : "+v"(rr));
        if constexpr ((VAR & VAR_SEL) != 0) asm volatile("" : "+v"(lob0), "+v"(lobW));
        const uint32_t ob0 = lob0, obW = lobW;
        // with one wave phase (z = 32, 64) the masks are constants; opaque,
        // they stay per-layer scalar loads rather than registers held (and,
        // in the larger kernels, spilled) across the iteration loop
        cmask_t* mtl = (cmask_t*)&WrapMasks<RATE, Z>::table.m[wphase][L][0];
        if constexpr ((VAR & VAR_SMASK) != 0 && WrapMasks<RATE, Z>::P == 1) asm volatile("" : "+s"(mtl));
        auto atl = [&](auto Lc2, auto Jc2) {
            if constexpr ((VAR & VAR_SEL) != 0 && ((VAR & VAR_SMASK) != 0 || SPLIT)) {
                return lds_ptr(sel_addr(Lc2, Jc2, Wc, ob0, obW, mtl));
            } else if constexpr ((VAR & VAR_SEL) != 0) {
                constexpr int b = CD::blk(L, decltype(Jc2)::value);
                constexpr int s = CD::sh(L, decltype(Jc2)::value);
                const uint32_t o = s == 0 ? ob0 : (rr >= Z - s ? obW : ob0);
                return lds_ptr(o + (uint32_t)((b * Z + s) * 4));
            } else if constexpr ((VAR & VAR_NOHOIST) != 0) {
                return at_r(Lc2, Jc2, rr);
            } else {
                return at(Lc2, Jc2);
            }
        };
        float x[D];
        float mn1 = 10000.0f, mn2 = 10000.0f;
        float a0 = 0.0f, a1 = 0.0f;   // |L| of the first two slots of the current triple
        uint32_t sx = 0, sprev = 0;
        static_for<0, D>([&](auto Jc) {
            constexpr int J = decltype(Jc)::value;
            const float v = *atl(Lc, Jc) - msg[L][J];
            if constexpr ((VAR & VAR_REREAD) == 0) x[J] = v;
            const float av = fabsf(v);
            // min1 / min2 = the two smallest |L| of the row (a tie at the
            // minimum gives min2 = min1, as the reference's rescan that skips
            // only the first argmin does), clamped at 10000.  Whole triples:
            // min3 / med3 give their two smallest, merged into the running
            // pair in three operations; the last D % 3 slots go one at a time.
            constexpr int Dt = D - D % 3;
            if constexpr (J < Dt) {
                if constexpr (J % 3 == 0) {
                    a0 = av;
                } else if constexpr (J % 3 == 1) {
                    a1 = av;
                } else {
                    const float lo = fminf(fminf(a0, a1), av);
                    const float mid = __builtin_amdgcn_fmed3f(a0, a1, av);
                    if constexpr (J == 2) {
                        mn2 = fminf(mn2, mid);
                    } else {
                        mn2 = fminf(fminf(mn2, fmaxf(mn1, lo)), mid);
                    }
                    mn1 = fminf(mn1, lo);
                }
            } else {
                mn2 = __builtin_amdgcn_fmed3f(mn1, mn2, av);
                mn1 = fminf(mn1, av);
            }
            // sign parity, two slots per three-input XOR
            if constexpr ((J & 1) == 0) {
                sprev = __float_as_uint(v);
                if constexpr (J == D - 1) sx ^= sprev;
            } else {
                sx = __builtin_amdgcn_bitop3_b32(sx, sprev, __float_as_uint(v), 0x96);   // a ^ b ^ c
            }
        });
        // message sign bit of slot n = signbit(L_n) ^ parity(signbits) ^ (D & 1)
        const uint32_t kw = (sx ^ ((uint32_t)(D & 1) << 31)) & 0x80000000u;
        if constexpr (MODE == MODE_NMS) {
            // s*(min*0.815) in double: the per-row magnitudes carry the parity
            // part of the sign; slot n flips it by signbit(L_n).  The stored
            // message is that double rounded once to float (:220), the
            // posterior is (double)L + it rounded once (:222-223).
            const double d1 = __dmul_rn((double)mn1, 0.815);
            const double d2 = __dmul_rn((double)mn2, 0.815);
            const uint32_t h1 = (uint32_t)__double2hiint(d1) | kw, h2 = (uint32_t)__double2hiint(d2) | kw;
            const int l1 = __double2loint(d1), l2 = __double2loint(d2);
            if constexpr ((VAR & VAR_REREAD) != 0) asm volatile("" ::: "memory");
            static_for<0, D>([&](auto Jc) {
                constexpr int J = decltype(Jc)::value;
                float v;
                if constexpr ((VAR & VAR_REREAD) != 0) v = *atl(Lc, Jc) - msg[L][J];
                else v = x[J];
                const bool e = fabsf(v) == mn1;
                const uint32_t hi = (e ? h2 : h1) ^ (__float_as_uint(v) & 0x80000000u);
                const double dsg = __hiloint2double((int)hi, e ? l2 : l1);
                msg[L][J] = __double2float_rn(dsg);
                *atl(Lc, Jc) = __double2float_rn(__dadd_rn((double)v, dsg));
            });
        } else if constexpr (MODE == MODE_ALT) {
            if constexpr ((VAR & VAR_REREAD) != 0) asm volatile("" ::: "memory");
            if (a.rule == 0) {
                // OMS (:225-231): max((double)min - 0.10, 0.0) as std::max
                // does, the row's sign parts as above, one rounding each
                const double t1 = __dsub_rn((double)mn1, 0.10), t2 = __dsub_rn((double)mn2, 0.10);
                const double d1 = t1 < 0.0 ? 0.0 : t1, d2 = t2 < 0.0 ? 0.0 : t2;
                const uint32_t h1 = (uint32_t)__double2hiint(d1) | kw, h2 = (uint32_t)__double2hiint(d2) | kw;
                const int l1 = __double2loint(d1), l2 = __double2loint(d2);
                static_for<0, D>([&](auto Jc) {
                    constexpr int J = decltype(Jc)::value;
                    float v;
                    if constexpr ((VAR & VAR_REREAD) != 0) v = *atl(Lc, Jc) - msg[L][J];
                    else v = x[J];
                    const bool e = fabsf(v) == mn1;
                    const uint32_t hi = (e ? h2 : h1) ^ (__float_as_uint(v) & 0x80000000u);
                    const double dsg = __hiloint2double((int)hi, e ? l2 : l1);
                    msg[L][J] = __double2float_rn(dsg);
                    *atl(Lc, Jc) = __double2float_rn(__dadd_rn((double)v, dsg));
                });
            } else {
                // MS (:234-237): s * min and L + s * min in float
                static_for<0, D>([&](auto Jc) {
                    constexpr int J = decltype(Jc)::value;
                    float v;
                    if constexpr ((VAR & VAR_REREAD) != 0) v = *atl(Lc, Jc) - msg[L][J];
                    else v = x[J];
                    const bool e = fabsf(v) == mn1;
                    const uint32_t sg = (__float_as_uint(v) ^ kw) & 0x80000000u;
                    const float m = __uint_as_float(__float_as_uint(e ? mn2 : mn1) | sg);
                    msg[L][J] = m;
                    *atl(Lc, Jc) = __fadd_rn(v, m);
                });
            }
        } else {
            const int tt = t < a.n_offset_iters ? t : a.n_offset_iters - 1;
            // the layer's offsets are uniform and read-only during the
            // launch: scalar loads through the constant address space (as
            // global loads they were re-issued around the volatile LDS
            // accesses: 1.20 against 0.77 ms per 65536-frame decode)
            const __attribute__((address_space(4))) float* beta =
                (const __attribute__((address_space(4))) float*)(a.offsets + ((size_t)tt * NL + L) * C);
            if constexpr ((VAR & VAR_REREAD) != 0) asm volatile("" ::: "memory");
            static_for<0, D>([&](auto Jc) {
                constexpr int J = decltype(Jc)::value;
                float v;
                if constexpr ((VAR & VAR_REREAD) != 0) v = *atl(Lc, Jc) - msg[L][J];
                else v = x[J];
                const bool e = fabsf(v) == mn1;
                const uint32_t sg = (__float_as_uint(v) ^ kw) & 0x80000000u;
                float q = (e ? mn2 : mn1) - beta[J];
                q = q > 0.0f ? q : 0.0f;
                const float m = __uint_as_float(__float_as_uint(q) | sg);
                msg[L][J] = m;
                *atl(Lc, Jc) = v + m;
            });
        }
    };

    // Whether this lane's row of layer L fails its parity check
    // (lib/ldpc_decoder.cc:67-104 over hd = rx >= 0), addressed like the
    // layers (one select per wrapped slot).  Each slot's compare is ballotted
    // into a wave-wide mask and the masks are XORed in scalar registers in
    // slot order; the barrier on the running mask keeps each read next to its
    // compare, so the check adds no live registers beside the messages.
    auto row_fails = [&](auto Lc, auto Wc) -> bool {
        constexpr int L = decltype(Lc)::value;
        int rr = r;
        if constexpr (uses_rr) asm volatile("" : "+v"(rr));
        const uint32_t ob0 = lds_addr(smem) + (uint32_t)(kLdsPad + ((active ? f : 0) * NS + r) * 4);
        const uint32_t obW = ob0 - (uint32_t)(Z * 4);
        cmask_t* mt = (cmask_t*)&WrapMasks<RATE, Z>::table.m[wphase][L][0];
        if constexpr ((VAR & VAR_SMASK) != 0) asm volatile("" : "+s"(mt));
        uint64_t pm = 0;
        static_for<0, CD::deg(L)>([&](auto Jc) {
            constexpr int J = decltype(Jc)::value;
            float xv;
            if constexpr ((VAR & VAR_SEL) != 0 && ((VAR & VAR_SMASK) != 0 || SPLIT)) {
                xv = *lds_ptr(sel_addr(Lc, Jc, Wc, ob0, obW, mt));
            } else if constexpr ((VAR & VAR_NOHOIST) != 0) {
                xv = *at_r(Lc, Jc, rr);
            } else {
                xv = *at(Lc, Jc);
            }
            pm ^= __builtin_amdgcn_ballot_w64(xv >= 0.0f);
            if constexpr ((J & 3) == 3) asm volatile("" : "+s"(pm));
        });
        return __builtin_amdgcn_inverse_ballot_w64(pm);
    };
    // Failing parity checks among this lane's rows: the syndrome weight.
    auto fail_rows = [&](auto Wc) -> int {
        int fails = 0;
        static_for<0, NL>([&](auto Lc) { fails += row_fails(Lc, Wc) ? 1 : 0; });
        return fails;
    };
    // The stop decisions only need to know whether some row of a frame fails
    // (compute_syndrome(hd, true) returns at the first failing row,
    // lib/ldpc_decoder.cc:67-104).  Lanes with `live` set cnt[f] = 1 as soon
    // as their row of a layer fails; after layers 0, 1, 3 and 7 the wave
    // re-reads the flags of its frames and checks no further layers once every
    // live frame it holds rows of is known to fail.  A wave's own flag writes
    // are visible to its later reads (LDS executes a wave's accesses in
    // order); other waves' may or may not be yet, which only decides how soon
    // it stops, not the result: the flags are read for the decision after a
    // barrier.  Frames close to convergence are checked in full; frames with
    // many errors fail in the first layer.
    auto mark_fails = [&](auto Wc, bool live) {
        bool need = __builtin_amdgcn_ballot_w64(live) != 0;
        static_for<0, NL>([&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            if (need) {
                if (row_fails(Lc, Wc) && live) cnt[f] = 1;
                if constexpr (L + 1 < NL && (L == 0 || L == 1 || L == 3 || L == 7))
                    need = __builtin_amdgcn_ballot_w64(live && cnt[f] == 0) != 0;
            }
        });
    };

    bool done = !active;
    int it_done = 0;
    // Initial check, iterations and final syndrome.  W: this wave's index
    // with a wave-split shape (each wave runs its own copy; every copy passes the same
    // barriers), -1 otherwise.
    auto run = [&](auto Wc) {
        if (a.early_stop) {
            // initial hard decision + syndrome (lib/ldpc_decoder.cc:130-139)
            mark_fails(Wc, active);
            __syncthreads();
            if (active && cnt[f] == 0) done = true;
            __syncthreads();
            if (tid < F) cnt[tid] = 0;
        }
        int any = __syncthreads_or(!done);
        // lanes still decoding, as a wave-wide mask in scalar registers: the
        // layer guard is then an exec-mask operation with no vector work
        uint64_t live = __builtin_amdgcn_ballot_w64(!done);
        for (int it = 0; it < a.max_iter && any; it++) {
            static_for<0, NL>([&](auto Lc) {
                if (__builtin_amdgcn_inverse_ballot_w64(live)) layer(Lc, it, Wc);
                __syncthreads();
            });
            if (a.early_stop) {
                // lib/ldpc_decoder.cc:242-251
                mark_fails(Wc, !done);
                __syncthreads();
                if (!done && cnt[f] == 0) {
                    done = true;
                    it_done = it + 1;
                }
                live = __builtin_amdgcn_ballot_w64(!done);
                any = __syncthreads_or(!done);
                if (tid < F) cnt[tid] = 0;
            }
        }
        __syncthreads();
        // Final syndrome: the weight where it is asked for; with fixed
        // iterations the return value only needs to know whether a row
        // fails; with early stop it follows from the stop decisions.  Frames
        // that stopped early satisfy every check (weight 0, cnt[f] stays 0).
        if (a.synd) {
            if (!done) {
                const int fr = fail_rows(Wc);
                if (fr) atomicAdd(&cnt[f], (uint32_t)fr);
            }
        } else if (a.ret && !a.early_stop) {
            mark_fails(Wc, !done);
        }
    };
    if constexpr (SPLIT) {
        const int w = __builtin_amdgcn_readfirstlane(tid) >> 6;
        static_for<0, Z / RS>([&](auto Wc) {
            if (w == decltype(Wc)::value) run(Wc);
        });
    } else {
        run(std::integral_constant<int, -1>{});
    }
    const bool converged = a.early_stop && done && active;
    if (!converged) it_done = a.max_iter;

    __syncthreads();
    if (active && r == 0) {
        unsigned long long g = frame0 + f;
        if constexpr (EXT) {
            if (!contiguous) g = fid[f];
        }
        const uint32_t w = cnt[f];
        if (a.synd) a.synd[g] = (uint16_t)w;
        if (a.iters) a.iters[g] = (uint8_t)it_done;
        if (a.ret) {
            uint8_t rv;
            if (a.early_stop) rv = (converged || a.max_iter == 0) ? 0 : 1;  // :124,141,261
            else rv = w != 0;
            a.ret[g] = rv;
            if (fail_out) fail_out[f] = rv;
        }
    }

    if (a.post) {
        for (int s = 0; s < Feff; s++) {
            f32x4* dst = reinterpret_cast<f32x4*>(a.post + fo(s));
            const f32x4* src = reinterpret_cast<const f32x4*>(rxall + s * NS);
            for (int i = tid; i < N / 4; i += blockDim.x) __builtin_nontemporal_store(src[i], dst + i);
        }
    }
    if (a.hard) {
        // per frame: ballots of 64 consecutive variables, each the 8 bytes of
        // packed hard decisions of its chunk, stored by lane 0 as two dwords
        // (n / 8 = 3 z bytes per frame: rows are 4-byte aligned, and the last
        // chunk of a frame holds 64 or 32 variables).  The chunk loop is
        // uniform over the wave, so its indices and the frame's output
        // address are scalar values.
        constexpr int kChunks = (N + 63) / 64;
        static_assert(N % 32 == 0 && (N / 8) % 4 == 0, "packed rows of whole dwords");
        const int lane = tid & 63;
        const int nw = blockDim.x >> 6;
        const int w0 = __builtin_amdgcn_readfirstlane(tid >> 6);
        for (int q = w0; q < Feff * kChunks; q += nw) {
            const int s = q / kChunks, c = q - s * kChunks;
            const int v = 64 * c + lane;
            const bool bit = v < N && rxall[s * NS + v] >= 0.0f;
            const uint64_t m = __builtin_amdgcn_ballot_w64(bit);
            unsigned long long off = fo(s);
            if constexpr (EXT) off = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)off)) |
                                     ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(off >> 32)) << 32);
            uint32_t* dst = reinterpret_cast<uint32_t*>(a.hard + off / 8 + 8 * c);
            if (lane == 0) {
                dst[0] = (uint32_t)m;
                if (64 * c + 64 <= N) dst[1] = (uint32_t)(m >> 32);
            }
        }
    }
}

// Waves per SIMD the compiler must allow for: 4 for the plain kernels of
// the wave-split shapes (their register budget then drops to 128, with a few
// spills outside the layer loop; 2.4 % faster at rate 5/6 z=96 than the 3
// waves their natural 132-133 registers give) and for the learned +
// re-queue kernel of those shapes; 4 for the plain kernels of the
// register-lean shape (111-139 registers unconstrained; never slower at 4
// waves, 4-7 % faster for the high-rate codes at their best frame counts)
// and of VAR_OCC4 shapes; no constraint otherwise.
template <int MODE, int VAR, bool EXT>
constexpr int min_waves_per_simd()
{
    if (split_rows<VAR>() != 0 && (!EXT || MODE == MODE_LEARNED_FB)) return 4;
    if (!EXT && (VAR == VAR_LEAN || (VAR & VAR_OCC4) != 0)) return 4;
    return 1;
}

template <int RATE, int Z, int MODE, int VAR = 0, bool EXT = false>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(min_waves_per_simd<MODE, VAR, EXT>())))
void tdmp_kernel(DecodeArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int F = a.frames_per_wg;
    if constexpr (!EXT) {
        // plain launch: tile blockIdx.x of a contiguous frame range
        const unsigned long long frame0 = (unsigned long long)blockIdx.x * (unsigned long long)F;
        if (frame0 >= a.nframes) return;   // whole workgroup, before any barrier
        const unsigned long long left = a.nframes - frame0;
        tdmp_tile<RATE, Z, MODE, VAR, false>(a, smem, frame0, left < (unsigned long long)F ? (int)left : F);
        return;
    }
    unsigned long long nfr = a.nframes;
    if (a.bucket) {
        const unsigned long long c = a.bucket[1];
        nfr = c < nfr ? c : nfr;
    }
    // One call site of the tile body for both launch modes: the plain launch
    // decodes tile blockIdx.x once; the persistent one takes tiles from the
    // counter until the list is exhausted (every workgroup reaches the exit:
    // the counter only grows).
    unsigned long long* slot = reinterpret_cast<unsigned long long*>(smem + kTileSlot);
    if constexpr (MODE != MODE_LEARNED_FB) {
        bool first = true;
        for (;;) {
            unsigned long long tile;
            if (a.tile_counter) {
                if (threadIdx.x == 0) *slot = atomicAdd(a.tile_counter, 1ull);
                __syncthreads();
                tile = *slot;
                __syncthreads();
            } else {
                if (!first) break;
                tile = blockIdx.x;
            }
            first = false;
            const unsigned long long frame0 = tile * (unsigned long long)F;
            if (frame0 >= nfr) break;   // uniform over the workgroup
            const unsigned long long left = nfr - frame0;
            tdmp_tile<RATE, Z, MODE, VAR, true>(a, smem, frame0, left < (unsigned long long)F ? (int)left : F);
            __syncthreads();
        }
    } else {
        // learned tiles; failures go to this workgroup's queue (LDS), decoded
        // F at a time with normalised min-sum, the rest after the last tile
        uint32_t* queue = reinterpret_cast<uint32_t*>(smem + kQueueOff);   // [2F]
        uint8_t* fails = smem + kFailOff;                                   // [F]
        uint32_t* qn = reinterpret_cast<uint32_t*>(smem + kQueueCountOff);
        if (threadIdx.x == 0) *qn = 0;
        DecodeArgs b = a;   // the fallback pass
        b.max_iter = a.fallback_iters;
        b.early_stop = 1;
        for (;;) {
            if (threadIdx.x == 0) *slot = atomicAdd(a.tile_counter, 1ull);
            __syncthreads();
            const unsigned long long frame0 = *slot * (unsigned long long)F;
            const bool last = frame0 >= nfr;
            __syncthreads();
            if (!last) {
                const unsigned long long left = nfr - frame0;
                const int fe = left < (unsigned long long)F ? (int)left : F;
                // a contiguous batch takes the plain tile form (8-10 % faster
                // at rate 2/3A z=96: fewer scalar registers spilled), frame
                // lists the extended one
                if (!a.frame_idx && !a.bucket && !a.frame_off)
                    tdmp_tile<RATE, Z, MODE_LEARNED, VAR, false>(a, smem, frame0, fe, nullptr, fails);
                else
                    tdmp_tile<RATE, Z, MODE_LEARNED, VAR, true>(a, smem, frame0, fe, nullptr, fails);
                __syncthreads();
                if (threadIdx.x == 0) {
                    const unsigned long long s0 = a.bucket ? (unsigned long long)a.bucket[0] : 0ull;
                    uint32_t n = *qn;
                    for (int s = 0; s < fe; s++) {
                        const uint32_t g = a.frame_idx ? a.frame_idx[s0 + frame0 + s] : (uint32_t)(frame0 + s);
                        if (a.fallback) a.fallback[g] = fails[s];
                        if (fails[s]) queue[n++] = g;
                    }
                    *qn = n;
                }
                __syncthreads();
            }
            const int q = (int)*qn;
            if (q >= F || (last && q > 0)) {
                const int take = q < F ? q : F;
                tdmp_tile<RATE, Z, MODE_NMS, VAR, true>(b, smem, 0, take, queue, nullptr);
                __syncthreads();
                if (threadIdx.x == 0) {
                    for (int i = take; i < q; i++) queue[i - take] = queue[i];
                    *qn = (uint32_t)(q - take);
                }
                __syncthreads();
            }
            if (last && *qn == 0) break;   // *qn is read by all after the barrier above
        }
    }
}

// Workgroups of this kernel resident per CU at frames_per_wg F (0 on error).
template <int RATE, int Z, int MODE, int VAR = 0, bool EXT = false>
int occupancy_tdmp(int F)
{
    if (split_rows<VAR>() != 0 && F != 64 / split_rows<VAR>()) return 0;
    const int threads = ((F * Z + 63) / 64) * 64;
    const size_t lds = kLdsPad + (size_t)F * lds_frame_stride<Code<RATE, Z>::N, VAR>() * sizeof(float);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tdmp_kernel<RATE, Z, MODE, VAR, EXT>, threads, lds) !=
        hipSuccess)
        return 0;
    return nb;
}

// A plain (contiguous, one tile per workgroup) launch of a shape compiled in
// the plain form only.
template <int RATE, int Z, int MODE, int VAR>
hipError_t launch_tdmp_plain(const DecodeArgs& a, hipStream_t stream)
{
    const int F = a.frames_per_wg;
    if (F < 1 || F > max_frames_per_wg<RATE, Z>()) return hipErrorInvalidValue;
    if (a.frame_idx || a.bucket || a.frame_off || a.tile_counter) return hipErrorInvalidValue;
    if (a.nframes == 0) return hipSuccess;
    const int threads = ((F * Z + 63) / 64) * 64;
    const size_t lds = kLdsPad + (size_t)F * lds_frame_stride<Code<RATE, Z>::N, VAR>() * sizeof(float);
    const unsigned long long grid = (a.nframes + F - 1) / F;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL((tdmp_kernel<RATE, Z, MODE, VAR, false>), dim3((unsigned)grid), dim3(threads), lds, stream, a);
    return hipGetLastError();
}

// EXT kernels serve indirect, ragged and persistent launches; the plain one
// contiguous batches.  The launcher picks by the arguments.
template <int RATE, int Z, int MODE, int VAR = 0>
hipError_t launch_tdmp(const DecodeArgs& a, hipStream_t stream)
{
    const int F = a.frames_per_wg;
    if (F < 1 || F > max_frames_per_wg<RATE, Z>()) return hipErrorInvalidValue;
    if (split_rows<VAR>() != 0 && F != 64 / split_rows<VAR>()) return hipErrorInvalidValue;   // wave-split lane mapping
    if (a.nframes == 0) return hipSuccess;
    const int threads = ((F * Z + 63) / 64) * 64;
    const size_t lds = kLdsPad + (size_t)F * lds_frame_stride<Code<RATE, Z>::N, VAR>() * sizeof(float);
    unsigned long long grid = (a.nframes + F - 1) / F;
    if (a.tile_counter && a.persistent_grid > 0 && (unsigned long long)a.persistent_grid < grid)
        grid = (unsigned long long)a.persistent_grid;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    // the learned + re-queue kernel is persistent: extended form only
    constexpr bool ext_only = MODE == MODE_LEARNED_FB;
    const bool ext = ext_only || a.frame_idx || a.bucket || a.frame_off || a.tile_counter;
    if (ext)
        hipLaunchKernelGGL((tdmp_kernel<RATE, Z, MODE, VAR, true>), dim3((unsigned)grid), dim3(threads), lds, stream, a);
    else if constexpr (!ext_only)
        hipLaunchKernelGGL((tdmp_kernel<RATE, Z, MODE, VAR, false>), dim3((unsigned)grid), dim3(threads), lds, stream, a);
    return hipGetLastError();
}

}  // namespace wl
