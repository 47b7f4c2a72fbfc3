// mixed.hip -- batches of frames with mixed (rate, z) in one captured launch.
//
// A frame's configuration id is rate * 19 + (z - 24) / 4 (114 codes).  Frame g
// holds n_g = 24 z_g LLRs starting at element frame_off[g] of a ragged buffer.
// One decode is a hipGraph, captured once per set of buffers and replayed:
//   prep      zero the per-code counters and persistent tile counters
//   level     (early stop) a reliability level per frame from the mean |LLR|
//             of its first 256 LLRs, the frames per level and per code in the
//             same pass, and a counting sort of the frame ids by level (scan /
//             scatter below, over 256 levels)
//   hist      (no level pass) frames per code (workgroup histogram in LDS,
//             then one atomic per code per workgroup); invalid ids counted
//             separately
//   scan      bucket start of each code (exclusive scan, one workgroup)
//   scatter   frame ids into their code's bucket, in level order: a bucket is
//             a sequence of short runs of frames of similar reliability, so the
//             frames of one workgroup tile tend to stop after similar numbers
//             of iterations (a workgroup runs until its slowest frame stops)
//   decode    one persistent kernel per lifting size z (group.hpp; graph
//             siblings on 4 streams) taking tiles of the buckets of all
//             rates at that z from one atomic counter, each rate with its own
//             code shape; an empty group exits at once.  (The learned graph
//             runs a per-code learned + re-queue kernel for each code with
//             offsets.)
// Per-frame results land at the frame's own index/offset, so the output does
// not depend on bucket order and equals a per-code decode bit for bit.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/wimax_ldpc_amd.h"
#include "channel_dev.hpp"
#include "dispatch.hpp"
#include "encoder.hpp"
#include "errors.hpp"
#include "group.hpp"

namespace wl {

constexpr int kCodes = kNumRates * kNumZ;   // 114
constexpr int kSideStreams = 4;

constexpr int kLevels = 256;
constexpr int kLevelSample = 256;   // LLRs per frame in the reliability estimate (n >= 576)

__global__ void mixed_prep_kernel(uint32_t* counts, uint32_t* cursors, uint32_t* lv_counts, uint32_t* lv_cursors,
                                  unsigned long long* tiles)
{
    for (int i = threadIdx.x; i <= kCodes; i += blockDim.x) {
        counts[i] = 0;
        cursors[i] = 0;
    }
    for (int i = threadIdx.x; i <= kLevels; i += blockDim.x) {
        lv_counts[i] = 0;
        lv_cursors[i] = 0;
    }
    for (int i = threadIdx.x; i < kCodes + kNumZ; i += blockDim.x) tiles[i] = 0;
}

// fallback[g] = 0 for every frame of the learned graph: the re-queue kernels
// then write 0 / 1 for their codes' frames, and the frames decoded by the
// group kernels (normalised min-sum only) keep 0
__global__ void zero_bytes_kernel(uint8_t* __restrict__ p, unsigned long long n)
{
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Reliability level of frame g: the mean |LLR| of its first kLevelSample LLRs
// as float bits >> 19 (exponent and 4 mantissa bits: 16 steps per octave),
// over 2^-8 .. 2^8; lower means, zero and NaN give 0, higher ones 255.  One
// wave per frame.  Only the order of frames within a code bucket depends on it.
// The same pass counts the frames per level (lv_counts) and per code (counts;
// codes >= kCodes in bin kCodes), as bucket_hist_kernel would over the level
// and code bytes: workgroup histograms in LDS, one atomic per non-empty bin.
constexpr int kLevelWaves = 16;   // waves per workgroup of the reliability pass (fewer histogram flushes)

__global__ __launch_bounds__(64 * kLevelWaves) void mixed_level_kernel(const uint8_t* __restrict__ cfg,
                                                          const unsigned long long* __restrict__ off,
                                                          const float* __restrict__ llr, unsigned long long n,
                                                          uint8_t* __restrict__ level, uint32_t* __restrict__ lv_counts,
                                                          uint32_t* __restrict__ counts)
{
    __shared__ uint32_t hl[kLevels];
    __shared__ uint32_t hc[kCodes + 1];
    for (int i = threadIdx.x; i < kLevels; i += blockDim.x) hl[i] = 0;
    for (int i = threadIdx.x; i <= kCodes; i += blockDim.x) hc[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // the wave's frame index is uniform: scalar loads of cfg and off
    const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (unsigned long long g = (unsigned long long)blockIdx.x * kLevelWaves + wave; g < n;
         g += (unsigned long long)gridDim.x * kLevelWaves) {
        const int c = cfg[g];
        float sum = 0.0f;
        if (c < kCodes) {
            const float* x = llr + off[g];
#pragma unroll
            for (int q = 0; q < kLevelSample / 64; q++) sum += fabsf(x[q * 64 + lane]);
        }
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (lane == 0) {
            const float mean = sum * (1.0f / kLevelSample);
            int lv = 0;
            if (mean > 0.0f) {
                const int k = (int)(__float_as_uint(mean) >> 19) - (127 - 8) * 16;
                lv = k < 0 ? 0 : (k >= kLevels ? kLevels - 1 : k);
            }
            level[g] = (uint8_t)lv;
            atomicAdd(&hl[lv], 1u);
            atomicAdd(&hc[c < kCodes ? c : kCodes], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLevels; i += blockDim.x)
        if (hl[i]) atomicAdd(&lv_counts[i], hl[i]);
    for (int i = threadIdx.x; i <= kCodes; i += blockDim.x)
        if (hc[i]) atomicAdd(&counts[i], hc[i]);
}

// Counting sort by a byte key with K buckets (keys >= K: counted in bin K,
// not scattered).  hist: per-key counts (workgroup histogram in LDS, then one
// atomic per key per workgroup).
template <int K>
__global__ __launch_bounds__(256) void bucket_hist_kernel(const uint8_t* __restrict__ key, unsigned long long n,
                                                          uint32_t* __restrict__ counts)
{
    __shared__ uint32_t h[K + 1];
    for (int i = threadIdx.x; i <= K; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (unsigned long long g = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
         g += (unsigned long long)gridDim.x * blockDim.x) {
        const int c = key[g];
        atomicAdd(&h[c < K ? c : K], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= K; i += blockDim.x)
        if (h[i]) atomicAdd(&counts[i], h[i]);
}

// buckets[k] = {start, count}.  One wave: lane l holds keys
// l * per .. l * per + per - 1; a shuffle scan over the lanes' sums gives
// each lane its first start.
template <int K>
__global__ __launch_bounds__(64) void bucket_scan_kernel(const uint32_t* __restrict__ counts,
                                                         uint32_t* __restrict__ buckets)
{
    constexpr int per = (K + 63) / 64;
    const int lane = threadIdx.x;
    uint32_t v[per];
    uint32_t local = 0;
#pragma unroll
    for (int j = 0; j < per; j++) {
        const int c = lane * per + j;
        v[j] = c < K ? counts[c] : 0u;
        local += v[j];
    }
    uint32_t incl = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    uint32_t s = incl - local;
#pragma unroll
    for (int j = 0; j < per; j++) {
        const int c = lane * per + j;
        if (c < K) {
            buckets[2 * c] = s;
            buckets[2 * c + 1] = v[j];
        }
        s += v[j];
    }
}

// Element i of the input order (order[i], or i without one) goes to its key's
// bucket.  Each workgroup takes a contiguous chunk of kScatterPer x 256
// elements: ranks within the chunk from LDS atomics, then one global
// reservation per key and workgroup (global atomics on the cursors from every
// element contend heavily).  Order inside a chunk's run is arbitrary; frames
// decode independently, so results do not depend on it.
constexpr int kScatterPer = 8;
template <int K>
__global__ __launch_bounds__(256) void bucket_scatter_kernel(const uint8_t* __restrict__ key,
                                                             const uint32_t* __restrict__ order,
                                                             unsigned long long n,
                                                             const uint32_t* __restrict__ buckets,
                                                             uint32_t* __restrict__ cursors,
                                                             uint32_t* __restrict__ out)
{
    __shared__ uint32_t h[K];
    __shared__ uint32_t base[K];
    for (int i = threadIdx.x; i < K; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const unsigned long long i0 = (unsigned long long)blockIdx.x * (256 * kScatterPer);
    int code[kScatterPer];
    uint32_t rank[kScatterPer], gid[kScatterPer];
#pragma unroll
    for (int q = 0; q < kScatterPer; q++) {
        const unsigned long long i = i0 + (unsigned long long)q * 256 + threadIdx.x;
        gid[q] = i < n ? (order ? order[i] : (uint32_t)i) : 0u;
        code[q] = i < n ? (int)key[gid[q]] : K;
        rank[q] = code[q] < K ? atomicAdd(&h[code[q]], 1u) : 0u;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += blockDim.x)
        base[i] = h[i] ? buckets[2 * i] + atomicAdd(&cursors[i], h[i]) : 0u;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kScatterPer; q++)
        if (code[q] < K) out[base[code[q]] + rank[q]] = gid[q];
}

// frame_off = exclusive prefix sum of n(cfg[g]); single pass per workgroup of
// 1024 frames plus a carry chain over workgroups (three kernels)
__device__ __forceinline__ unsigned long long code_n(int c) { return c < kCodes ? 24ull * (24 + 4 * (c % kNumZ)) : 0ull; }

__global__ __launch_bounds__(1024) void off_block_kernel(const uint8_t* __restrict__ cfg, unsigned long long n,
                                                         unsigned long long* __restrict__ off,
                                                         unsigned long long* __restrict__ block_sum)
{
    __shared__ unsigned long long s[1024];
    const unsigned long long g = (unsigned long long)blockIdx.x * 1024 + threadIdx.x;
    const unsigned long long v = g < n ? code_n(cfg[g]) : 0ull;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {   // Hillis-Steele inclusive scan
        const unsigned long long t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0ull;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (g < n) off[g] = s[threadIdx.x] - v;
    if (threadIdx.x == 1023) block_sum[blockIdx.x] = s[1023];
}

__global__ void off_carry_kernel(unsigned long long* block_sum, unsigned long long nblocks)
{
    if (threadIdx.x == 0) {
        unsigned long long acc = 0;
        for (unsigned long long b = 0; b < nblocks; b++) {
            const unsigned long long v = block_sum[b];
            block_sum[b] = acc;
            acc += v;
        }
        block_sum[nblocks] = acc;
    }
}

__global__ __launch_bounds__(1024) void off_add_kernel(unsigned long long* __restrict__ off, unsigned long long n,
                                                       const unsigned long long* __restrict__ block_sum)
{
    const unsigned long long g = (unsigned long long)blockIdx.x * 1024 + threadIdx.x;
    if (g < n) off[g] += block_sum[blockIdx.x];
}

// Mixed channel: one frame per workgroup, the frame's own code; frame g of
// code c equals frame g of wl_channel_dev on an encoder of code c.
__global__ __launch_bounds__(128) void mixed_channel_kernel(const EncodeDesc* __restrict__ descs,
                                                            const uint8_t* __restrict__ cfg,
                                                            const unsigned long long* __restrict__ off,
                                                            unsigned long long nframes,
                                                            unsigned long long seed, float eb_lo, float eb_hi,
                                                            float* __restrict__ llr, uint8_t* __restrict__ bits,
                                                            float* __restrict__ ebn0_out)
{
    __shared__ uint8_t c[2304 + 96];
    for (unsigned long long g = blockIdx.x; g < nframes; g += gridDim.x) {
        const int code = cfg[g];
        if (code >= kCodes) continue;   // uniform per workgroup
        generate_frame(descs[code], seed, g, eb_lo, eb_hi, c, llr + off[g], bits ? bits + off[g] : nullptr,
                       ebn0_out ? ebn0_out + g : nullptr);
        __syncthreads();
    }
}

}  // namespace wl

using namespace wl;

struct wl_mixed {
    int max_iter = 10;
    unsigned flags = 0;
    int num_cu = 0;
    const DecodeEntry* entries[kCodes] = {};
    int F[kCodes] = {};
    bool enabled[kCodes] = {};
    uint32_t *counts = nullptr, *cursors = nullptr, *buckets = nullptr, *idx = nullptr;
    uint32_t *lv_counts = nullptr, *lv_cursors = nullptr, *lv_buckets = nullptr, *lv_idx = nullptr;
    uint8_t* level = nullptr;
    unsigned long long* tiles = nullptr;
    size_t idx_cap = 0;
    unsigned long long* block_sum = nullptr;
    size_t bs_cap = 0;
    EncodeDesc* descs = nullptr;
    hipStream_t side[kSideStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kSideStreams] = {};
    hipStream_t cap_stream = nullptr;
    // captured graphs: [0] plain decode, [1] learned + fallback
    hipGraph_t graph[2] = {};
    hipGraphExec_t gexec[2] = {};
    const void* key[2][10] = {};   // buffers each graph refers to
    size_t key_n[2] = {};
    int key_it[2] = {};
    // learned offsets per rate: [T][mb][c]
    float* offsets[kNumRates] = {};
    int T[kNumRates] = {};
    int ncu = 0;
    int gocc[kNumZ] = {};   // resident workgroups per CU of each z group kernel
    int gF[kNumZ] = {};     // frames per workgroup of each z group kernel
};

namespace {

int mfail(int code, const char* msg)
{
    set_error_text(msg);
    return code;
}

int hipf(hipError_t e, const char* what)
{
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    set_error_text(buf);
    return WL_EDEVICE;
}

void drop_graph(wl_mixed* m)
{
    for (int i = 0; i < 2; i++) {
        if (m->gexec[i]) (void)hipGraphExecDestroy(m->gexec[i]);
        if (m->graph[i]) (void)hipGraphDestroy(m->graph[i]);
        m->gexec[i] = nullptr;
        m->graph[i] = nullptr;
    }
}

int ensure_idx(wl_mixed* m, size_t n)
{
    if (n <= m->idx_cap) return WL_OK;
    void* ps[] = {m->idx, m->lv_idx, m->level};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    m->idx = m->lv_idx = nullptr;
    m->level = nullptr;
    m->idx_cap = 0;
    const size_t nn = std::max<size_t>(n, 1);
    hipError_t e = hipMalloc((void**)&m->idx, nn * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->lv_idx, nn * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->level, nn);
    if (e != hipSuccess) return hipf(e, "hipMalloc");
    m->idx_cap = n;
    drop_graph(m);
    return WL_OK;
}

}  // namespace

extern "C" {

int wl_mixed_create(int max_iter, unsigned flags, wl_mixed** out)
{
    if (!out) return mfail(WL_EINVAL, "null output");
    *out = nullptr;
    if (max_iter < 0 || max_iter > 255) return mfail(WL_EINVAL, "max_iter out of range 0..255");
    if (flags & (WL_FLAG_OMS | WL_FLAG_MS)) return mfail(WL_ENOTSUP, "mixed batches decode with normalised min-sum");
    int dev = 0;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return hipf(e, "device query");
    wl_mixed* m = new (std::nothrow) wl_mixed();
    if (!m) return mfail(WL_ENOMEM, "out of host memory");
    m->max_iter = max_iter;
    m->flags = flags;
    m->num_cu = prop.multiProcessorCount;
    std::vector<EncodeDesc> descs(kCodes);
    for (int c = 0; c < kCodes; c++) {
        const int rate = c / kNumZ, z = z_of_index(c % kNumZ);
        m->entries[c] = decode_entry(rate, z);
        int f = 0, v = 0;
        if (wl_tuned_launch(rate, z, &f, &v) < 0) {
            delete m;
            return WL_EINVAL;
        }
        m->F[c] = f;   // the learned + re-queue kernels use the code's own tuned shape
        (void)v;
        m->enabled[c] = true;
        descs[c] = make_encode_desc_for(rate, z);
    }
    for (int zi = 0; zi < kNumZ; zi++) {
        const GroupEntry* ge = group_entry(z_of_index(zi));
        m->gF[zi] = ge->frames_per_wg;
        m->gocc[zi] = ge->occ(m->gF[zi]);
    }
    e = hipMalloc((void**)&m->counts, (kCodes + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->cursors, (kCodes + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->buckets, 2 * kCodes * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->lv_counts, (kLevels + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->lv_cursors, (kLevels + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->lv_buckets, 2 * kLevels * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->tiles, (kCodes + kNumZ) * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&m->descs, kCodes * sizeof(EncodeDesc));
    if (e == hipSuccess) e = hipMemcpy(m->descs, descs.data(), kCodes * sizeof(EncodeDesc), hipMemcpyHostToDevice);
    for (int i = 0; i < kSideStreams && e == hipSuccess; i++) {
        e = hipStreamCreateWithFlags(&m->side[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev_join[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        wl_mixed_destroy(m);
        return hipf(e, "mixed decoder allocation");
    }
    *out = m;
    return WL_OK;
}

int wl_mixed_destroy(wl_mixed* m)
{
    if (!m) return WL_OK;
    drop_graph(m);
    void* ps[] = {m->counts, m->cursors, m->buckets, m->idx, m->tiles, m->block_sum, m->descs,
                  m->lv_counts, m->lv_cursors, m->lv_buckets, m->lv_idx, m->level};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (int r = 0; r < kNumRates; r++)
        if (m->offsets[r]) (void)hipFree(m->offsets[r]);
    for (int i = 0; i < kSideStreams; i++) {
        if (m->side[i]) (void)hipStreamDestroy(m->side[i]);
        if (m->ev_join[i]) (void)hipEventDestroy(m->ev_join[i]);
    }
    if (m->ev_fork) (void)hipEventDestroy(m->ev_fork);
    if (m->cap_stream) (void)hipStreamDestroy(m->cap_stream);
    delete m;
    return WL_OK;
}

int wl_mixed_set_codes(wl_mixed* m, const int* codes, int ncodes)
{
    if (!m) return mfail(WL_EINVAL, "null mixed decoder");
    if (ncodes < 0 || (ncodes > 0 && !codes)) return mfail(WL_EINVAL, "bad code list");
    bool en[kCodes] = {};
    for (int i = 0; i < ncodes; i++) {
        if (codes[i] < 0 || codes[i] >= kCodes) return mfail(WL_EINVAL, "code id out of range 0..113");
        en[codes[i]] = true;
    }
    for (int c = 0; c < kCodes; c++) m->enabled[c] = ncodes == 0 ? true : en[c];
    drop_graph(m);
    return WL_OK;
}

int wl_mixed_get_launch(const wl_mixed* m, int rate, int z, int* frames_per_wg, int* variant, int* grid)
{
    if (!m || !frames_per_wg || !variant || !grid) return mfail(WL_EINVAL, "null argument");
    if (rate < 0 || rate >= kNumRates || !valid_z(z)) return mfail(WL_EINVAL, "invalid rate/z");
    const int zi = index_of_z(z);
    *frames_per_wg = m->gF[zi];
    *variant = 1;
    *grid = std::max(1, m->gocc[zi]) * m->num_cu;
    return WL_OK;
}

int wl_mixed_set_group_frames_per_wg(wl_mixed* m, int z, int frames_per_wg)
{
    if (!m) return mfail(WL_EINVAL, "null mixed decoder");
    if (!valid_z(z)) return mfail(WL_EINVAL, "invalid z");
    const GroupEntry* ge = group_entry(z);
    const int f = frames_per_wg == 0 ? ge->frames_per_wg : frames_per_wg;
    if (f < 1 || f > ge->fmax) return mfail(WL_EINVAL, "frames_per_wg out of range for this z");
    const int occ = ge->occ(f);
    if (occ < 1) return mfail(WL_EINVAL, "no workgroup of this shape fits a CU");
    const int zi = index_of_z(z);
    m->gF[zi] = f;
    m->gocc[zi] = occ;
    drop_graph(m);
    return WL_OK;
}

int wl_mixed_offsets_dev(wl_mixed* m, const uint8_t* cfg, size_t nframes, uint64_t* frame_off, void* stream)
{
    if (!m || !cfg || !frame_off) return mfail(WL_EINVAL, "null argument");
    if (nframes == 0) return WL_OK;
    const size_t nb = (nframes + 1023) / 1024;
    if (nb + 1 > m->bs_cap) {
        if (m->block_sum) (void)hipFree(m->block_sum);
        m->block_sum = nullptr;
        m->bs_cap = 0;
        hipError_t e = hipMalloc((void**)&m->block_sum, (nb + 1) * sizeof(unsigned long long));
        if (e != hipSuccess) return hipf(e, "hipMalloc");
        m->bs_cap = nb + 1;
    }
    hipStream_t s = (hipStream_t)stream;
    auto* off = reinterpret_cast<unsigned long long*>(frame_off);
    hipLaunchKernelGGL(off_block_kernel, dim3((unsigned)nb), dim3(1024), 0, s, cfg, (unsigned long long)nframes,
                       off, m->block_sum);
    hipLaunchKernelGGL(off_carry_kernel, dim3(1), dim3(64), 0, s, m->block_sum, (unsigned long long)nb);
    hipLaunchKernelGGL(off_add_kernel, dim3((unsigned)nb), dim3(1024), 0, s, off, (unsigned long long)nframes,
                       m->block_sum);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WL_OK : hipf(e, "offset kernels");
}

int wl_mixed_channel_dev(wl_mixed* m, uint64_t seed, const uint8_t* cfg, const uint64_t* frame_off,
                         size_t nframes, float ebn0_lo, float ebn0_hi, float* llr, uint8_t* bits, float* ebn0,
                         void* stream)
{
    if (!m || !cfg || !frame_off || !llr) return mfail(WL_EINVAL, "null argument");
    if (nframes == 0) return WL_OK;
    const unsigned grid = (unsigned)std::min<size_t>(nframes, (size_t)m->num_cu * 16);
    hipLaunchKernelGGL(mixed_channel_kernel, dim3(grid), dim3(128), 0, (hipStream_t)stream, m->descs, cfg,
                       reinterpret_cast<const unsigned long long*>(frame_off), (unsigned long long)nframes,
                       (unsigned long long)seed, ebn0_lo, ebn0_hi, llr, bits, ebn0);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WL_OK : hipf(e, "mixed channel kernel");
}

static int capture_decode(wl_mixed* m, int learned, int fallback_iters, const uint8_t* cfg,
                          const uint64_t* frame_off, const float* llr, size_t nframes, float* post, uint8_t* hard,
                          uint8_t* ret, uint8_t* iters, uint16_t* synd, uint8_t* fallback)
{
    hipStream_t s = m->cap_stream;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hipf(e, "begin capture");
    const unsigned long long n = nframes;
    const unsigned blocks = (unsigned)std::min<unsigned long long>((n + 255) / 256, (unsigned long long)m->num_cu * 8);
    const unsigned sblocks = (unsigned)((n + 256 * kScatterPer - 1) / (256 * kScatterPer));
    hipLaunchKernelGGL(mixed_prep_kernel, dim3(1), dim3(256), 0, s, m->counts, m->cursors, m->lv_counts,
                       m->lv_cursors, m->tiles);
    if (learned && fallback)
        hipLaunchKernelGGL(zero_bytes_kernel, dim3(blocks), dim3(256), 0, s, fallback, n);
    // Reliability order: only where frames stop early (with fixed iterations
    // every frame costs the same).
    const bool ordered = (learned || (m->flags & WL_FLAG_EARLY_STOP)) && !(m->flags & WL_FLAG_MIXED_UNORDERED);
    if (ordered) {
        const unsigned lblocks = (unsigned)std::min<unsigned long long>((n + kLevelWaves - 1) / kLevelWaves,
                                                                        (unsigned long long)m->num_cu * 4);
        // levels and both histograms in one pass
        hipLaunchKernelGGL(mixed_level_kernel, dim3(lblocks), dim3(64 * kLevelWaves), 0, s, cfg,
                           reinterpret_cast<const unsigned long long*>(frame_off), llr, n, m->level, m->lv_counts,
                           m->counts);
        hipLaunchKernelGGL(bucket_scan_kernel<kLevels>, dim3(1), dim3(64), 0, s, m->lv_counts, m->lv_buckets);
        hipLaunchKernelGGL(bucket_scatter_kernel<kLevels>, dim3(sblocks), dim3(256), 0, s, m->level,
                           (const uint32_t*)nullptr, n, m->lv_buckets, m->lv_cursors, m->lv_idx);
    } else {
        hipLaunchKernelGGL(bucket_hist_kernel<kCodes>, dim3(blocks), dim3(256), 0, s, cfg, n, m->counts);
    }
    hipLaunchKernelGGL(bucket_scan_kernel<kCodes>, dim3(1), dim3(64), 0, s, m->counts, m->buckets);
    hipLaunchKernelGGL(bucket_scatter_kernel<kCodes>, dim3(sblocks), dim3(256), 0, s, cfg,
                       ordered ? (const uint32_t*)m->lv_idx : (const uint32_t*)nullptr, n, m->buckets, m->cursors,
                       m->idx);
    e = hipEventRecord(m->ev_fork, s);
    for (int i = 0; i < kSideStreams && e == hipSuccess; i++) e = hipStreamWaitEvent(m->side[i], m->ev_fork, 0);
    // Jobs: the learned + re-queue kernel of each code that has offsets (in
    // the learned graph), and one group kernel per z for every other enabled
    // code of that z.  Longest first onto the least-loaded side stream, by
    // edges per frame (z x nonzero base blocks) summed over the job's codes:
    // per-code frame counts are only known on the device, so the estimate
    // takes the codes of the batch as equally likely.
    auto work = [](int c) {
        const LayerTable lt = make_layer_table(c / kNumZ);
        int nnz = 0;
        for (int l = 0; l < lt.nl; l++) nnz += lt.deg[l];
        return (long)(24 + 4 * (c % kNumZ)) * nnz;
    };
    struct Job {
        int code;        // >= 0: per-code learned kernel; -1: group of z index zi
        int zi;
        unsigned mask;   // group: rates decoded
        long w;
    };
    Job jobs[kCodes + kNumZ];
    int njobs = 0;
    for (int zi = 0; zi < kNumZ; zi++) {
        Job gj{-1, zi, 0u, 0};
        for (int rate = 0; rate < kNumRates; rate++) {
            const int c = rate * kNumZ + zi;
            if (!m->enabled[c]) continue;
            if (learned && m->offsets[rate] && m->entries[c]->learned_fb) {
                jobs[njobs++] = Job{c, zi, 0u, work(c)};
            } else {
                gj.mask |= 1u << rate;
                gj.w += work(c);
            }
        }
        if (gj.mask) jobs[njobs++] = gj;
    }
    std::stable_sort(jobs, jobs + njobs, [](const Job& x, const Job& y) { return x.w > y.w; });
    DecodeArgs base{};
    base.llr = llr;
    base.post = post;
    base.hard = hard;
    base.ret = ret;
    base.iters = iters;
    base.synd = synd;
    base.frame_idx = m->idx;
    base.frame_off = reinterpret_cast<const unsigned long long*>(frame_off);
    base.nframes = n;
    base.max_iter = m->max_iter;
    base.early_stop = (m->flags & WL_FLAG_EARLY_STOP) ? 1 : 0;
    long load[kSideStreams] = {};
    for (int ji = 0; ji < njobs && e == hipSuccess; ji++) {
        const Job& j = jobs[ji];
        const int si = (int)(std::min_element(load, load + kSideStreams) - load);
        load[si] += j.w;
        hipStream_t ss = m->side[si];
        if (j.code < 0) {
            GroupArgs g{};
            g.a = base;
            g.buckets = m->buckets;
            g.tile_counter = m->tiles + kCodes + j.zi;
            g.enabled = j.mask;
            g.frames_per_wg = m->gF[j.zi];
            const int grid = std::max(1, m->gocc[j.zi]) * m->num_cu;
            e = group_entry(z_of_index(j.zi))->launch(g, grid, ss);
        } else {
            const int c = j.code, rate = c / kNumZ;
            DecodeArgs a = base;
            a.bucket = m->buckets + 2 * c;
            a.tile_counter = m->tiles + c;
            a.frames_per_wg = m->F[c];
            a.offsets = m->offsets[rate];
            a.n_offset_iters = m->T[rate];
            a.fallback_iters = fallback_iters;
            a.fallback = fallback;
            const int occ = m->entries[c]->occ_learned_fb(m->F[c]);
            a.persistent_grid = (int)std::min<unsigned long long>((unsigned long long)std::max(1, occ) * m->num_cu,
                                                                  (n + m->F[c] - 1) / m->F[c]);
            e = m->entries[c]->learned_fb(a, ss);
        }
    }
    for (int i = 0; i < kSideStreams && e == hipSuccess; i++) {
        e = hipEventRecord(m->ev_join[i], m->side[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, m->ev_join[i], 0);
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hipf(e != hipSuccess ? e : e2, "mixed decode capture");
    }
    e = hipGraphInstantiate(&m->gexec[learned], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hipf(e, "graph instantiate");
    }
    m->graph[learned] = g;
    return WL_OK;
}

static int run_graph(wl_mixed* m, int learned, int fallback_iters, const uint8_t* cfg, const uint64_t* frame_off,
                     const float* llr, size_t nframes, float* post, uint8_t* hard, uint8_t* ret, uint8_t* iters,
                     uint16_t* synd, uint8_t* fallback, hipStream_t stream)
{
    int rc = ensure_idx(m, nframes);
    if (rc) return rc;
    const void* key[10] = {cfg, frame_off, llr, post, hard, ret, iters, synd, fallback, nullptr};
    if (!m->gexec[learned] || m->key_n[learned] != nframes || m->key_it[learned] != fallback_iters ||
        memcmp(key, m->key[learned], sizeof key) != 0) {
        if (m->gexec[learned]) (void)hipGraphExecDestroy(m->gexec[learned]);
        if (m->graph[learned]) (void)hipGraphDestroy(m->graph[learned]);
        m->gexec[learned] = nullptr;
        m->graph[learned] = nullptr;
        rc = capture_decode(m, learned, fallback_iters, cfg, frame_off, llr, nframes, post, hard, ret, iters, synd,
                            fallback);
        if (rc) return rc;
        memcpy(m->key[learned], key, sizeof key);
        m->key_n[learned] = nframes;
        m->key_it[learned] = fallback_iters;
    }
    hipError_t e = hipGraphLaunch(m->gexec[learned], stream);
    return e == hipSuccess ? WL_OK : hipf(e, "graph launch");
}

int wl_mixed_decode_dev(wl_mixed* m, const uint8_t* cfg, const uint64_t* frame_off, const float* llr,
                        size_t nframes, float* post, uint8_t* hard, uint8_t* ret, uint8_t* iters, uint16_t* synd,
                        void* stream)
{
    if (!m) return mfail(WL_EINVAL, "null mixed decoder");
    if (nframes == 0) return WL_OK;
    if (!cfg || !frame_off || !llr) return mfail(WL_EINVAL, "cfg, frame_off and llr are required");
    if (nframes > 0xffffffffull) return mfail(WL_EINVAL, "nframes too large");
    return run_graph(m, 0, 0, cfg, frame_off, llr, nframes, post, hard, ret, iters, synd, nullptr,
                     (hipStream_t)stream);
}

int wl_mixed_set_offsets(wl_mixed* m, int rate, const float* offsets_host, int T)
{
    if (!m) return mfail(WL_EINVAL, "null mixed decoder");
    if (rate < 0 || rate >= kNumRates) return mfail(WL_EINVAL, "invalid rate");
    if (m->offsets[rate]) (void)hipFree(m->offsets[rate]);
    m->offsets[rate] = nullptr;
    m->T[rate] = 0;
    drop_graph(m);
    if (!offsets_host) return WL_OK;
    if (T < 1 || T > 255) return mfail(WL_EINVAL, "T out of range 1..255");
    const LayerTable lt = make_layer_table(rate);
    const size_t np = (size_t)T * lt.nl * lt.c;
    hipError_t e = hipMalloc((void**)&m->offsets[rate], np * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(m->offsets[rate], offsets_host, np * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hipf(e, "offsets upload");
    m->T[rate] = T;
    return WL_OK;
}

int wl_mixed_decode_learned_fallback_dev(wl_mixed* m, const uint8_t* cfg, const uint64_t* frame_off,
                                         const float* llr, size_t nframes, int fallback_iters, float* post,
                                         uint8_t* hard, uint8_t* ret, uint8_t* iters, uint16_t* synd,
                                         uint8_t* fallback, void* stream)
{
    if (!m) return mfail(WL_EINVAL, "null mixed decoder");
    if (nframes == 0) return WL_OK;
    if (!cfg || !frame_off || !llr) return mfail(WL_EINVAL, "cfg, frame_off and llr are required");
    if (nframes > 0xffffffffull) return mfail(WL_EINVAL, "nframes too large");
    if (fallback_iters < 0 || fallback_iters > 255) return mfail(WL_EINVAL, "fallback_iters out of range");
    if (post && (const float*)post == llr)
        return mfail(WL_EINVAL, "post is llr: re-queued frames are decoded again from the input LLRs");
    return run_graph(m, 1, fallback_iters, cfg, frame_off, llr, nframes, post, hard, ret, iters, synd, fallback,
                     (hipStream_t)stream);
}

int wl_mixed_bucket_list(wl_mixed* m, uint32_t* buckets_host, uint32_t* idx_dev, size_t nframes)
{
    if (!m || !buckets_host) return mfail(WL_EINVAL, "null argument");
    if (idx_dev && nframes > m->idx_cap) return mfail(WL_EINVAL, "nframes exceeds the last decode's batch");
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(buckets_host, m->buckets, 2 * kCodes * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && idx_dev && nframes)
        e = hipMemcpy(idx_dev, m->idx, nframes * sizeof(uint32_t), hipMemcpyDeviceToDevice);
    if (e != hipSuccess) return hipf(e, "read bucket list");
    return WL_OK;
}

int wl_mixed_invalid_count(wl_mixed* m, uint64_t* out)
{
    if (!m || !out) return mfail(WL_EINVAL, "null argument");
    uint32_t v[kCodes + 1] = {};
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(v, m->counts, sizeof v, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipf(e, "read frame counts");
    uint64_t skipped = v[kCodes];
    for (int c = 0; c < kCodes; c++)
        if (!m->enabled[c]) skipped += v[c];
    *out = skipped;
    return WL_OK;
}

}  // extern "C"
