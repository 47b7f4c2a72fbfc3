// capi.hip -- the C ABI (include/wimax_ldpc_amd.h) over the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/wimax_ldpc_amd.h"
#include "dispatch.hpp"
#include "encoder.hpp"
#include "errors.hpp"
#include "generic.hpp"
#include "qc_host.hpp"
#include "tdmp_pair.hpp"
#include "tuned.hpp"

#define WL_VERSION_STR "wimax_ldpc_amd 0.1 (gfx950)"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(WL_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define WL_HIP(call)                                   \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

struct HostCode {
    wl_code_info info{};
    wl::LayerTable lt{};
    std::vector<int16_t> table;   // [m][c]
};

int make_code(int rate, int z, HostCode& hc)
{
    if (rate < 0 || rate >= wl::kNumRates)
        return fail(WL_EINVAL, "invalid coderate %d (valid: 0..5)", rate);
    if (!wl::valid_z(z)) return fail(WL_EINVAL, "invalid z factor %d (valid: 24..96 step 4)", z);
    hc.lt = wl::make_layer_table(rate);
    const wl_base_info& bi = wl_base_table[rate];
    wl_code_info& ci = hc.info;
    ci.rate = rate;
    ci.z = z;
    ci.n = 24 * z;
    ci.mb = hc.lt.nl;
    ci.m = hc.lt.nl * z;
    ci.k = ci.n - ci.m;
    ci.c = hc.lt.c;
    ci.d = wl::lift(rate, bi.d0, z);
    ci.x = bi.x;
    hc.table.assign((size_t)ci.m * ci.c, (int16_t)-1);
    for (int l = 0; l < hc.lt.nl; l++)
        for (int r = 0; r < z; r++)
            for (int j = 0; j < hc.lt.deg[l]; j++) {
                const int s = wl::lift(rate, hc.lt.p96[l][j], z);
                hc.table[(size_t)(l * z + r) * ci.c + j] = (int16_t)(hc.lt.blk[l][j] * z + (r + s) % z);
            }
    return WL_OK;
}

wl::EncodeDesc make_encode_desc(const HostCode& hc) { return wl::make_encode_desc_for(hc.info.rate, hc.info.z); }

// Default frames per workgroup: the most lanes busy, then the larger group up
// to 256 threads (the autotuner may override per (rate, z)).
int default_frames_per_wg(int z, int fmax)
{
    int best = 1;
    double best_eff = -1.0;
    for (int f = 1; f <= fmax; f++) {
        const int thr = ((f * z + 63) / 64) * 64;
        if (thr > 512) break;
        const double eff = (double)(f * z) / thr;
        const bool better = eff > best_eff + 1e-9 || (eff > best_eff - 1e-9 && thr <= 256);
        if (better) {
            best = f;
            best_eff = eff;
        }
    }
    return best;
}

bool tuned_lookup(int rate, int z, int* f, int* variant)
{
    for (const wl::TunedShape& t : wl::kTunedShapes)
        if (t.rate == rate && t.z == z) {
            *f = t.f;
            *variant = t.variant;
            return true;
        }
    return false;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    int ensure(size_t n)
    {
        if (n <= cap) return WL_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc");
        cap = n;
        return WL_OK;
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Pinned host memory (hipHostMalloc: mapped into the device's address
// space) that kernels read and write directly.  The host calls of small
// batches stage through it: a call is then a host copy in, one kernel and a
// synchronisation, with no DMA transfer of pageable memory.
template <class T>
struct PinnedBuf {
    T* p = nullptr;
    size_t cap = 0;
    int ensure(size_t n)
    {
        if (n <= cap) return WL_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
        cap = n;
        return WL_OK;
    }
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};

// Host calls up to this many bytes of input go through pinned memory the
// kernels access over PCIe; larger ones are copied to device buffers first.
constexpr size_t kPinnedCallBytes = 4u << 20;

// the handle's stream for host calls, created on first use
int host_stream(hipStream_t* s)
{
    if (*s) return WL_OK;
    hipError_t e = hipStreamCreateWithFlags(s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        *s = nullptr;
        return hip_fail(e, "hipStreamCreate");
    }
    return WL_OK;
}

// fallback list header {first slot 0, count 0}
__global__ void zero_bucket_kernel(uint32_t* p)
{
    p[0] = 0;
    p[1] = 0;
}

// Frames with ret != 0 go to the fallback list; fallback[g] records it.
__global__ void compact_failed_kernel(const uint8_t* __restrict__ ret, unsigned long long n,
                                      uint32_t* __restrict__ idx, uint32_t* __restrict__ count,
                                      uint8_t* __restrict__ fallback)
{
    const unsigned long long g = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool bad = g < n && ret[g] != 0;
    if (g < n && fallback) fallback[g] = bad ? 1 : 0;
    const unsigned long long m = __ballot(bad);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, 0);
    if (bad) idx[base + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)g;
}

}  // namespace

void wl::set_error_text(const char* text) { g_err = text; }

struct wl_decoder {
    HostCode code;
    int max_iter = 10;
    unsigned flags = 0;
    int frames_per_wg = 0;
    int variant = 0;
    const wl::DecodeEntry* entry = nullptr;
    // small-batch kernel (contiguous NMS batches of at most small_batch frames)
    wl::PairLaunch pair = nullptr;
    size_t small_batch = 0;
    DevBuf<float> offsets;
    int T = 0;
    // host-call staging
    DevBuf<float> d_llr;
    DevBuf<uint8_t> d_hard, d_ret, d_iters;
    DevBuf<int8_t> d_dec;   // 0/1 bytes of the host-call decisions
    // run-time quasi-cyclic code (wl_decoder_create_qc): descriptor on the device
    bool generic = false;
    wl::GenDesc gen{};
    DevBuf<wl::GenDesc> d_gen;
    // fallback scratch
    DevBuf<uint32_t> d_idx, d_count;
    // small host calls: posteriors in place in pinned memory, then status
    PinnedBuf<float> h_llr;
    PinnedBuf<uint8_t> h_stat;
    hipStream_t stream = nullptr;   // host calls
    ~wl_decoder()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// packed hard decisions [nframes][n/8] (bit v%8 of byte v/8) -> one 0/1 byte
// per variable, the `decoded` format of ldpc_decoder::decode; four output
// bytes per thread
__global__ void unpack_hard_kernel(const uint8_t* __restrict__ hard, int8_t* __restrict__ out,
                                   unsigned long long nbytes4)
{
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes4; i += stride) {
        // output bytes 4i .. 4i+3 are bits 4i%8 .. of packed byte 4i/8
        const uint32_t b = hard[i >> 1] >> ((i & 1) * 4);
        const uint32_t w = (b & 1u) | ((b >> 1) & 1u) << 8 | ((b >> 2) & 1u) << 16 | ((b >> 3) & 1u) << 24;
        reinterpret_cast<uint32_t*>(out)[i] = w;
    }
}

// the same for any n: [nframes][ceil(n/8)] -> [nframes][n], one byte per thread
__global__ void unpack_hard_any_kernel(const uint8_t* __restrict__ hard, int8_t* __restrict__ out, int n,
                                       unsigned long long total)
{
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    const unsigned long long nb8 = (unsigned long long)(n + 7) / 8;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const unsigned long long f = i / (unsigned long long)n, v = i - f * n;
        out[i] = (int8_t)((hard[f * nb8 + v / 8] >> (v % 8)) & 1u);
    }
}

struct wl_encoder {
    HostCode code;
    wl::EncodeDesc desc{};
    DevBuf<uint8_t> d_info, d_cw;
    PinnedBuf<uint8_t> h_info, h_cw;   // small host calls
    hipStream_t stream = nullptr;      // host calls
    ~wl_encoder()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

const char* wl_last_error(void) { return g_err.c_str(); }

const char* wl_version(void) { return WL_VERSION_STR; }

int wl_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int wl_code_params(int rate, int z, wl_code_info* out)
{
    if (!out) return fail(WL_EINVAL, "null output");
    HostCode hc;
    int rc = make_code(rate, z, hc);
    if (rc) return rc;
    *out = hc.info;
    return WL_OK;
}

int wl_code_table(int rate, int z, int16_t* out, size_t out_len)
{
    HostCode hc;
    int rc = make_code(rate, z, hc);
    if (rc) return rc;
    if (!out || out_len < hc.table.size())
        return fail(WL_EINVAL, "table needs %zu entries", hc.table.size());
    memcpy(out, hc.table.data(), hc.table.size() * sizeof(int16_t));
    return WL_OK;
}

// the small-batch threshold for the current device (256 compute units where
// none can be queried)
static size_t default_small_batch(int z, bool early_stop)
{
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) {
        (void)hipGetLastError();
        ncu = 256;
    }
    return wl::pair_default_max_frames(z, ncu, early_stop);
}

int wl_decoder_create(int rate, int z, int max_iter, unsigned flags, wl_decoder** out)
{
    if (!out) return fail(WL_EINVAL, "null output");
    *out = nullptr;
    if (max_iter < 0 || max_iter > 255) return fail(WL_EINVAL, "max_iter %d out of range 0..255", max_iter);
    if ((flags & WL_FLAG_OMS) && (flags & WL_FLAG_MS)) return fail(WL_EINVAL, "WL_FLAG_OMS and WL_FLAG_MS together");
    wl_decoder* d = new (std::nothrow) wl_decoder();
    if (!d) return fail(WL_ENOMEM, "out of host memory");
    int rc = make_code(rate, z, d->code);
    if (rc) {
        delete d;
        return rc;
    }
    d->max_iter = max_iter;
    d->flags = flags;
    d->entry = wl::decode_entry(rate, z);
    int tf = 0, tv = 0;
    if (tuned_lookup(rate, z, &tf, &tv) && tf >= 1 && tf <= d->entry->fmax && tv >= 0 && tv < wl::kNumVariants &&
        d->entry->nms[tv] && (wl::variant_frames(tv) == 0 || tf == wl::variant_frames(tv))) {
        d->frames_per_wg = tf;
        d->variant = tv;
    } else {
        d->frames_per_wg = default_frames_per_wg(z, d->entry->fmax);
        d->variant = 0;
    }
    d->pair = wl::pair_launcher(rate, z);
    d->small_batch = default_small_batch(z, (flags & WL_FLAG_EARLY_STOP) != 0);
    *out = d;
    return WL_OK;
}

int wl_decoder_set_small_batch(wl_decoder* dec, size_t max_frames)
{
    if (!dec) return fail(WL_EINVAL, "null decoder");
    if (dec->generic || !dec->pair) return max_frames == 0 || max_frames == WL_SMALL_BATCH_AUTO
                                               ? WL_OK
                                               : fail(WL_ENOTSUP, "run-time codes have no small-batch kernel");
    dec->small_batch = max_frames == WL_SMALL_BATCH_AUTO
                           ? default_small_batch(dec->code.info.z, (dec->flags & WL_FLAG_EARLY_STOP) != 0)
                           : max_frames;
    return WL_OK;
}

int wl_decoder_get_small_batch(const wl_decoder* dec, size_t* max_frames)
{
    if (!dec || !max_frames) return fail(WL_EINVAL, "null argument");
    // offset / plain min-sum and learned decoders never take the small-batch kernel
    const bool nms = !(dec->flags & (WL_FLAG_OMS | WL_FLAG_MS)) && dec->T == 0;
    *max_frames = dec->pair && nms ? dec->small_batch : 0;
    return WL_OK;
}

int wl_decoder_destroy(wl_decoder* dec)
{
    delete dec;
    return WL_OK;
}

int wl_tuned_launch(int rate, int z, int* frames_per_wg, int* variant)
{
    if (!frames_per_wg || !variant) return fail(WL_EINVAL, "null output");
    HostCode hc;
    int rc = make_code(rate, z, hc);
    if (rc) return rc;
    if (!tuned_lookup(rate, z, frames_per_wg, variant)) {
        *frames_per_wg = default_frames_per_wg(z, wl::decode_entry(rate, z)->fmax);
        *variant = 0;
        return 1;
    }
    return WL_OK;
}

int wl_decoder_set_variant(wl_decoder* dec, int variant)
{
    if (!dec) return fail(WL_EINVAL, "null decoder");
    if (dec->generic) return variant == 0 ? WL_OK : fail(WL_EINVAL, "run-time codes have one code shape");
    if (variant < 0 || variant >= wl::kNumVariants)
        return fail(WL_EINVAL, "variant %d out of range 0..%d", variant, wl::kNumVariants - 1);
    if (!dec->entry->nms[variant]) return fail(WL_EINVAL, "variant %d does not exist at this z", variant);
    const int vf = wl::variant_frames(variant);
    if (vf && dec->frames_per_wg != vf)
        return fail(WL_EINVAL, "variant %d decodes %d frames per workgroup (set frames_per_wg to %d first)", variant,
                    vf, vf);
    dec->variant = variant;
    return WL_OK;
}

int wl_decoder_get_variant(const wl_decoder* dec, int* out)
{
    if (!dec || !out) return fail(WL_EINVAL, "null argument");
    *out = dec->variant;
    return WL_OK;
}

int wl_decoder_set_frames_per_wg(wl_decoder* dec, int f)
{
    if (!dec) return fail(WL_EINVAL, "null decoder");
    if (dec->generic) return (f == 0 || f == 1) ? WL_OK : fail(WL_EINVAL, "run-time codes decode one frame per workgroup");
    if (f == 0) {
        // the autotuned launch shape: frames per workgroup and code shape
        int tf = 0, tv = 0;
        if (tuned_lookup(dec->code.info.rate, dec->code.info.z, &tf, &tv) && tf >= 1 && tf <= dec->entry->fmax &&
            tv >= 0 && tv < wl::kNumVariants && dec->entry->nms[tv] &&
            (wl::variant_frames(tv) == 0 || tf == wl::variant_frames(tv))) {
            dec->frames_per_wg = tf;
            dec->variant = tv;
        } else {
            dec->frames_per_wg = default_frames_per_wg(dec->code.info.z, dec->entry->fmax);
            dec->variant = 0;
        }
        return WL_OK;
    }
    if (f < 1 || f > dec->entry->fmax)
        return fail(WL_EINVAL, "frames_per_wg %d out of range 1..%d", f, dec->entry->fmax);
    const int vf = wl::variant_frames(dec->variant);
    if (vf && f != vf) dec->variant = 1;   // the same shape without the wave-split lane mapping
    dec->frames_per_wg = f;
    return WL_OK;
}

int wl_decoder_get_frames_per_wg(const wl_decoder* dec, int* out)
{
    if (!dec || !out) return fail(WL_EINVAL, "null argument");
    *out = dec->frames_per_wg;
    return WL_OK;
}

int wl_decoder_set_offsets(wl_decoder* dec, const float* offsets_host, int T)
{
    if (!dec) return fail(WL_EINVAL, "null decoder");
    if (dec->generic && offsets_host) return fail(WL_ENOTSUP, "learned offsets need a compiled 802.16e code");
    if (!offsets_host) {
        dec->T = 0;
        return WL_OK;
    }
    if (T < 1 || T > 255) return fail(WL_EINVAL, "T %d out of range 1..255", T);
    const size_t n = (size_t)T * dec->code.info.mb * dec->code.info.c;
    int rc = dec->offsets.ensure(n);
    if (rc) return rc;
    WL_HIP(hipMemcpy(dec->offsets.p, offsets_host, n * sizeof(float), hipMemcpyHostToDevice));
    dec->T = T;
    return WL_OK;
}

// learned: the learned offsets; otherwise the decoder's message rule, or
// normalised min-sum whatever the rule with force_nms (the learned fallback)
static int decode_launch(wl_decoder* dec, bool learned, int max_iter, bool early_stop,
                         const float* llr, size_t nframes, float* post, uint8_t* hard, uint8_t* ret,
                         uint8_t* iters, uint16_t* synd, const uint32_t* idx, const uint32_t* bucket,
                         hipStream_t s, bool force_nms = false)
{
    if (dec->generic) {
        if (learned || idx || bucket) return fail(WL_ENOTSUP, "run-time codes decode with normalised min-sum only");
        wl::GenArgs g{};
        g.desc = dec->d_gen.p;
        g.llr = llr;
        g.post = post;
        g.hard = hard;
        g.ret = ret;
        g.iters = iters;
        g.synd = synd;
        g.nframes = nframes;
        g.max_iter = max_iter;
        g.early_stop = early_stop ? 1 : 0;
        g.n = dec->gen.n;
        g.edges = dec->gen.edges;
        g.z = dec->gen.z;
        hipError_t e = wl::launch_generic(g, s);
        if (e != hipSuccess) return hip_fail(e, "generic decoder kernel launch");
        return WL_OK;
    }
    wl::DecodeArgs a{};
    a.llr = llr;
    a.post = post;
    a.hard = hard;
    a.ret = ret;
    a.iters = iters;
    a.synd = synd;
    a.offsets = dec->offsets.p;
    a.frame_idx = idx;
    a.bucket = bucket;
    a.nframes = nframes;
    a.frames_per_wg = dec->frames_per_wg;
    a.max_iter = max_iter;
    a.early_stop = early_stop ? 1 : 0;
    a.n_offset_iters = dec->T;
    hipError_t e;
    if (learned) {
        e = dec->entry->learned(a, s);
    } else if (!force_nms && (dec->flags & (WL_FLAG_OMS | WL_FLAG_MS))) {
        a.rule = (dec->flags & WL_FLAG_OMS) ? 0 : 1;
        e = dec->entry->alt(a, s);
    } else if (dec->pair && !idx && !bucket && nframes <= dec->small_batch) {
        e = dec->pair(a, s);
    } else {
        e = dec->entry->nms[dec->variant](a, s);
    }
    if (e != hipSuccess) return hip_fail(e, "decoder kernel launch");
    return WL_OK;
}

int wl_decode_f32_dev(wl_decoder* dec, const float* llr, size_t nframes, float* post,
                      uint8_t* hard, uint8_t* ret, uint8_t* iters, uint16_t* synd, void* stream)
{
    if (!dec) return fail(WL_EINVAL, "null decoder");
    if (nframes == 0) return WL_OK;
    if (!llr) return fail(WL_EINVAL, "null llr");
    if (nframes > 0xffffffffull) return fail(WL_EINVAL, "nframes %zu too large", nframes);
    const bool learned = dec->T > 0;
    return decode_launch(dec, learned, dec->max_iter, (dec->flags & WL_FLAG_EARLY_STOP) != 0, llr,
                         nframes, post, hard, ret, iters, synd, nullptr, nullptr, (hipStream_t)stream);
}

}  // extern "C"

// Scratch of wl_decode_learned_fallback_noalloc_dev: [0..1] the counter,
// [4 ..) the fallback list, in every device's copy of this array.  In host
// code the name is only a handle for hipGetSymbolAddress: kernels are given
// the address resolved on each device (device_scratch below).
static __device__ __align__(256) uint32_t scratch[4 + (size_t)WL_NOALLOC_MAX_FRAMES];

namespace {

// Argument checks shared by the two learned-fallback entry points.  The
// failed frames are decoded again from their input LLRs, so the posterior
// output may not overlap the input (in place would re-decode the learned
// posteriors instead).
int check_learned_fallback(const wl_decoder* dec, const float* llr, size_t nframes, int fallback_iters,
                           const float* post, const uint8_t* ret)
{
    if (!dec) return fail(WL_EINVAL, "null decoder");
    if (dec->T <= 0) return fail(WL_EINVAL, "no learned offsets set");
    if (dec->generic) return fail(WL_ENOTSUP, "learned offsets need a compiled 802.16e code");
    if (nframes == 0) return WL_OK;
    if (!llr || !ret) return fail(WL_EINVAL, "llr and ret are required");
    if (fallback_iters < 0 || fallback_iters > 255) return fail(WL_EINVAL, "fallback_iters out of range");
    if (nframes > 0xffffffffull) return fail(WL_EINVAL, "nframes %zu too large", nframes);
    if (post) {
        const size_t len = nframes * (size_t)dec->code.info.n;
        if (post < llr + len && llr < post + len)
            return fail(WL_EINVAL, "post overlaps llr: re-queued frames are decoded again from the input LLRs");
    }
    return WL_OK;
}

// The learned decode with fallback over caller-provided scratch: counter
// (two uint32 = the persistent kernel's 64-bit tile counter, or the
// fallback list header {first slot, count}) and idx [nframes] (the list of
// failed frames, used where z is not 24 or 96).
int learned_fallback(wl_decoder* dec, const float* llr, size_t nframes, int fallback_iters, float* post,
                     uint8_t* hard, uint8_t* ret, uint8_t* iters, uint16_t* synd, uint8_t* fallback,
                     uint32_t* counter, uint32_t* idx, hipStream_t s)
{
    hipLaunchKernelGGL(zero_bucket_kernel, dim3(1), dim3(1), 0, s, counter);
    if (dec->entry->learned_fb) {
        // one persistent kernel: learned tiles, failures re-queued in LDS and
        // decoded again with normalised min-sum by the same workgroup
        wl::DecodeArgs a{};
        a.llr = llr;
        a.post = post;
        a.hard = hard;
        a.ret = ret;
        a.iters = iters;
        a.synd = synd;
        a.offsets = dec->offsets.p;
        a.tile_counter = reinterpret_cast<unsigned long long*>(counter);
        a.nframes = nframes;
        a.frames_per_wg = dec->frames_per_wg;
        a.max_iter = dec->max_iter;
        a.early_stop = (dec->flags & WL_FLAG_EARLY_STOP) ? 1 : 0;
        a.n_offset_iters = dec->T;
        a.fallback_iters = fallback_iters;
        a.fallback = fallback;
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess) {
            const int occ = dec->entry->occ_learned_fb(dec->frames_per_wg);
            a.persistent_grid = (occ > 0 ? occ : 1) * ncu;
        }
        hipError_t e = dec->entry->learned_fb(a, s);
        if (e != hipSuccess) return hip_fail(e, "learned fallback kernel launch");
        return WL_OK;
    }
    int rc = decode_launch(dec, true, dec->max_iter, (dec->flags & WL_FLAG_EARLY_STOP) != 0, llr, nframes, post,
                           hard, ret, iters, synd, nullptr, nullptr, s);
    if (rc) return rc;
    const unsigned blocks = (unsigned)((nframes + 255) / 256);
    hipLaunchKernelGGL(compact_failed_kernel, dim3(blocks), dim3(256), 0, s, ret, (unsigned long long)nframes, idx,
                       counter + 1, fallback);
    WL_HIP(hipGetLastError());
    return decode_launch(dec, false, fallback_iters, true, llr, nframes, post, hard, ret, iters, synd, idx, counter,
                         s, true);
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[kMaxDevices] = {};

// this device's copy of scratch (resolved on its first use)
int device_scratch(uint32_t** out)
{
    int dev = 0;
    WL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(WL_ENOTSUP, "device %d beyond the %d the library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        WL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[dev];
    return WL_OK;
}

}  // namespace

extern "C" {

int wl_decode_learned_fallback_dev(wl_decoder* dec, const float* llr, size_t nframes,
                                   int fallback_iters, float* post, uint8_t* hard, uint8_t* ret,
                                   uint8_t* iters, uint16_t* synd, uint8_t* fallback, void* stream)
{
    int rc = check_learned_fallback(dec, llr, nframes, fallback_iters, post, ret);
    if (rc || nframes == 0) return rc;
    if ((rc = dec->d_count.ensure(2))) return rc;
    if (!dec->entry->learned_fb && (rc = dec->d_idx.ensure(nframes))) return rc;
    return learned_fallback(dec, llr, nframes, fallback_iters, post, hard, ret, iters, synd, fallback, dec->d_count.p,
                            dec->d_idx.p, (hipStream_t)stream);
}

int wl_decode_learned_fallback_noalloc_dev(wl_decoder* dec, const float* llr, size_t nframes,
                                           int fallback_iters, float* post, uint8_t* hard, uint8_t* ret,
                                           uint8_t* iters, uint16_t* synd, uint8_t* fallback, void* stream)
{
    int rc = check_learned_fallback(dec, llr, nframes, fallback_iters, post, ret);
    if (rc || nframes == 0) return rc;
    if (nframes > (size_t)WL_NOALLOC_MAX_FRAMES)
        return fail(WL_EINVAL, "nframes %zu exceeds the allocation-free limit of %d frames per call", nframes,
                    WL_NOALLOC_MAX_FRAMES);
    uint32_t* sc = nullptr;
    if ((rc = device_scratch(&sc))) return rc;
    return learned_fallback(dec, llr, nframes, fallback_iters, post, hard, ret, iters, synd, fallback, sc, sc + 4,
                            (hipStream_t)stream);
}

int wl_decode_f32_host(wl_decoder* dec, float* rx, size_t nframes, int8_t* decoded, uint8_t* ret,
                       uint8_t* iters)
{
    if (!dec) return fail(WL_EINVAL, "null decoder");
    if (nframes == 0) return WL_OK;
    if (!rx || !decoded || !ret) return fail(WL_EINVAL, "null argument");
    const int n = dec->code.info.n;
    const size_t nb = ((size_t)n + 7) / 8;
    int rc;
    if (nframes * n * sizeof(float) <= kPinnedCallBytes) {
        // small batch (decode() is one frame): the kernel reads the LLRs from
        // pinned memory and writes the posteriors back in place and the
        // status beside them; the 0/1 decisions are the posteriors' signs,
        // sliced here as the GPU slices them (rx >= 0, :130,242)
        const size_t len = nframes * (size_t)n;
        if ((rc = dec->h_llr.ensure(len)) || (rc = dec->h_stat.ensure(2 * nframes)) || (rc = host_stream(&dec->stream)))
            return rc;
        memcpy(dec->h_llr.p, rx, len * sizeof(float));
        uint8_t* hret = dec->h_stat.p;
        uint8_t* hit = dec->h_stat.p + nframes;
        rc = wl_decode_f32_dev(dec, dec->h_llr.p, nframes, dec->h_llr.p, nullptr, hret, hit, nullptr, dec->stream);
        if (rc) return rc;
        WL_HIP(hipStreamSynchronize(dec->stream));
        memcpy(rx, dec->h_llr.p, len * sizeof(float));
        for (size_t i = 0; i < len; i++) decoded[i] = rx[i] >= 0.0f ? 1 : 0;
        memcpy(ret, hret, nframes);
        if (iters) memcpy(iters, hit, nframes);
        return WL_OK;
    }
    if ((rc = dec->d_llr.ensure(nframes * n)) || (rc = dec->d_hard.ensure(nframes * nb)) ||
        (rc = dec->d_ret.ensure(nframes)) || (rc = dec->d_iters.ensure(nframes)) ||
        (rc = dec->d_dec.ensure(nframes * n)))
        return rc;
    WL_HIP(hipMemcpy(dec->d_llr.p, rx, nframes * n * sizeof(float), hipMemcpyHostToDevice));
    rc = wl_decode_f32_dev(dec, dec->d_llr.p, nframes, dec->d_llr.p /* in place, as decode() */,
                           dec->d_hard.p, dec->d_ret.p, dec->d_iters.p, nullptr, nullptr);
    if (rc) return rc;
    if (n % 8 == 0) {
        // whole packed bytes per frame (always for 802.16e: n = 24 z, z a multiple of 4)
        const unsigned long long n4 = (unsigned long long)nframes * n / 4;
        const unsigned blocks = (unsigned)std::min<unsigned long long>((n4 + 255) / 256, 65536ull);
        hipLaunchKernelGGL(unpack_hard_kernel, dim3(blocks), dim3(256), 0, nullptr, dec->d_hard.p, dec->d_dec.p, n4);
    } else {
        const unsigned long long total = (unsigned long long)nframes * n;
        const unsigned blocks = (unsigned)std::min<unsigned long long>((total + 255) / 256, 65536ull);
        hipLaunchKernelGGL(unpack_hard_any_kernel, dim3(blocks), dim3(256), 0, nullptr, dec->d_hard.p, dec->d_dec.p,
                           n, total);
    }
    WL_HIP(hipGetLastError());
    WL_HIP(hipMemcpy(rx, dec->d_llr.p, nframes * n * sizeof(float), hipMemcpyDeviceToHost));
    WL_HIP(hipMemcpy(decoded, dec->d_dec.p, nframes * n, hipMemcpyDeviceToHost));
    WL_HIP(hipMemcpy(ret, dec->d_ret.p, nframes, hipMemcpyDeviceToHost));
    if (iters) WL_HIP(hipMemcpy(iters, dec->d_iters.p, nframes, hipMemcpyDeviceToHost));
    return WL_OK;
}

int wl_decode_i8_host(wl_decoder* dec, const int8_t* rx, size_t nframes, int8_t* decoded, uint8_t* ret,
                      uint8_t* iters)
{
    if (!dec) return fail(WL_EINVAL, "null decoder");
    if (nframes == 0) return WL_OK;
    if (!rx) return fail(WL_EINVAL, "null rx");
    const int n = dec->code.info.n;
    std::vector<float> soft(nframes * n);
    for (size_t i = 0; i < soft.size(); i++) soft[i] = (float)(rx[i] * 2.0 - 1.0);  // :115
    return wl_decode_f32_host(dec, soft.data(), nframes, decoded, ret, iters);
}

int wl_compute_syndrome_host(const wl_decoder* dec, const int8_t* hd, int exit_early, unsigned* out)
{
    if (!dec || !hd || !out) return fail(WL_EINVAL, "null argument");
    const wl_code_info& ci = dec->code.info;
    unsigned errors = 0;
    for (int i = 0; i < ci.m; i++) {
        int8_t x = 0;
        for (int k = 0; k < ci.c; k++) {
            const int16_t v = dec->code.table[(size_t)i * ci.c + k];
            if (v != -1) x ^= hd[v];
        }
        if (x != 0) {
            if (exit_early) {
                errors = 1;
                break;
            }
            errors++;
        }
    }
    *out = errors;
    return WL_OK;
}

int wl_decoder_create_qc(const wl_qc_code* q, int max_iter, unsigned flags, wl_decoder** out)
{
    if (!out || !q) return fail(WL_EINVAL, "null argument");
    *out = nullptr;
    if (max_iter < 0 || max_iter > 255) return fail(WL_EINVAL, "max_iter %d out of range 0..255", max_iter);
    if (flags & (WL_FLAG_OMS | WL_FLAG_MS))
        return fail(WL_ENOTSUP, "run-time codes decode with normalised min-sum only");
    wl_decoder* d = new (std::nothrow) wl_decoder();
    if (!d) return fail(WL_ENOMEM, "out of host memory");
    d->code.info = q->info;
    d->code.table = q->table;
    d->max_iter = max_iter;
    d->flags = flags;
    d->generic = true;
    d->gen = q->desc;
    d->frames_per_wg = 1;
    d->variant = 0;
    int rc = d->d_gen.ensure(1);
    if (!rc) {
        hipError_t e = hipMemcpy(d->d_gen.p, &d->gen, sizeof(wl::GenDesc), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy");
    }
    if (rc) {
        delete d;
        return rc;
    }
    *out = d;
    return WL_OK;
}

int wl_encoder_create_qc(const wl_qc_code* q, wl_encoder** out)
{
    if (!out || !q) return fail(WL_EINVAL, "null argument");
    *out = nullptr;
    const wl_code_info& ci = q->info;
    if (ci.x < 0) return fail(WL_ENOTSUP, "the code has no dual-diagonal parity part");
    if (q->mb > wl::kMaxLayers || ci.c > wl::kMaxDeg || q->z > 127 || q->nb > 127 || (ci.n % 2) != 0)
        return fail(WL_ENOTSUP, "encoder limits: <= %d base rows, <= %d blocks per row, z <= 127, nb <= 127, even n",
                    wl::kMaxLayers, wl::kMaxDeg);
    const int F = std::max(1, 256 / q->z);
    if ((size_t)F * (ci.n + q->z) > 65536) return fail(WL_ENOTSUP, "frame too large for the encoder");
    wl_encoder* e = new (std::nothrow) wl_encoder();
    if (!e) return fail(WL_ENOMEM, "out of host memory");
    e->code.info = ci;
    e->code.table = q->table;
    wl::EncodeDesc& d = e->desc;
    d = wl::EncodeDesc{};
    d.z = q->z;
    d.n = ci.n;
    d.m = ci.m;
    d.k = ci.k;
    d.mb = q->mb;
    d.kb = q->nb - q->mb;
    d.x = ci.x;
    d.d = ci.d;
    d.frames_per_wg = F;
    for (int l = 0; l < q->mb; l++) {
        d.deg[l] = (int8_t)q->desc.deg[l];
        for (int j = 0; j < q->desc.deg[l]; j++) {
            d.blk[l][j] = (int8_t)q->desc.blk[l][j];
            d.sh[l][j] = (int8_t)q->desc.sh[l][j];
        }
    }
    *out = e;
    return WL_OK;
}

int wl_encoder_create(int rate, int z, wl_encoder** out)
{
    if (!out) return fail(WL_EINVAL, "null output");
    *out = nullptr;
    wl_encoder* e = new (std::nothrow) wl_encoder();
    if (!e) return fail(WL_ENOMEM, "out of host memory");
    int rc = make_code(rate, z, e->code);
    if (rc) {
        delete e;
        return rc;
    }
    e->desc = make_encode_desc(e->code);
    *out = e;
    return WL_OK;
}

int wl_encoder_create_standard(int rate, int z, wl_encoder** out)
{
    int rc = wl_encoder_create(rate, z, out);
    if (rc) return rc;
    if (!wl::make_standard_encode_desc(rate, z, &(*out)->desc)) {
        wl_encoder_destroy(*out);
        *out = nullptr;
        return fail(WL_ENOTSUP, "parity part of rate %d is not dual-diagonal", rate);
    }
    return WL_OK;
}

int wl_encoder_destroy(wl_encoder* enc)
{
    delete enc;
    return WL_OK;
}

int wl_encode_u8_dev(wl_encoder* enc, const uint8_t* info, size_t nframes, uint8_t* codeword, void* stream)
{
    if (!enc) return fail(WL_EINVAL, "null encoder");
    if (nframes == 0) return WL_OK;
    if (!info || !codeword) return fail(WL_EINVAL, "null buffer");
    if (nframes > 0xffffffffull) return fail(WL_EINVAL, "nframes %zu too large", nframes);
    hipError_t e = wl::launch_encode(enc->desc, info, codeword, nframes, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "encoder kernel launch");
    return WL_OK;
}

int wl_encode_u8_host(wl_encoder* enc, const uint8_t* info, size_t nframes, uint8_t* codeword)
{
    if (!enc) return fail(WL_EINVAL, "null encoder");
    if (nframes == 0) return WL_OK;
    if (!info || !codeword) return fail(WL_EINVAL, "null buffer");
    const size_t k = enc->code.info.k, n = enc->code.info.n;
    int rc;
    if (nframes * n <= kPinnedCallBytes) {
        // small batch (encode_data() is one frame): the kernel reads and
        // writes pinned memory
        if ((rc = enc->h_info.ensure(nframes * k)) || (rc = enc->h_cw.ensure(nframes * n)) ||
            (rc = host_stream(&enc->stream)))
            return rc;
        memcpy(enc->h_info.p, info, nframes * k);
        rc = wl_encode_u8_dev(enc, enc->h_info.p, nframes, enc->h_cw.p, enc->stream);
        if (rc) return rc;
        WL_HIP(hipStreamSynchronize(enc->stream));
        memcpy(codeword, enc->h_cw.p, nframes * n);
        return WL_OK;
    }
    if ((rc = enc->d_info.ensure(nframes * k)) || (rc = enc->d_cw.ensure(nframes * n))) return rc;
    WL_HIP(hipMemcpy(enc->d_info.p, info, nframes * k, hipMemcpyHostToDevice));
    rc = wl_encode_u8_dev(enc, enc->d_info.p, nframes, enc->d_cw.p, nullptr);
    if (rc) return rc;
    WL_HIP(hipDeviceSynchronize());
    WL_HIP(hipMemcpy(codeword, enc->d_cw.p, nframes * n, hipMemcpyDeviceToHost));
    return WL_OK;
}

int wl_channel_dev(wl_encoder* enc, uint64_t seed, uint64_t frame_offset, size_t nframes, float ebn0_lo,
                   float ebn0_hi, float* llr, uint8_t* bits, float* ebn0, void* stream)
{
    if (!enc) return fail(WL_EINVAL, "null encoder");
    if (nframes == 0) return WL_OK;
    if (nframes > 0xffffffffull) return fail(WL_EINVAL, "nframes %zu too large", nframes);
    hipError_t e = wl::launch_channel(enc->desc, seed, frame_offset, nframes, ebn0_lo, ebn0_hi, llr, bits,
                                      ebn0, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "channel kernel launch");
    return WL_OK;
}

}  // extern "C"
