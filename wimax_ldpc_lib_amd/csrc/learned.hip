// learned.hip -- training kernels of the learned offset min-sum decoder.
//
// Model: the layered decoder of tdmp.hpp unrolled for T iterations, where
// check-to-variable messages are  m = s * max(mag - beta[t][l][j], 0)  in FP32
// (one learned offset per nonzero base-matrix block (l, j) per iteration t),
// mag = min1, or min2 for the row's first strict argmin.  Loss: mean binary
// cross-entropy of the final posteriors over all n coded bits (bit 1 <=>
// posterior > 0, the reference's sign convention), or the weighted sum of
// that cross-entropy after every unrolled iteration (the gradient of an
// earlier iteration's term enters the reverse sweep at that iteration).
//
//   learned_fwd_kernel  forward over T iterations; per (frame, t, row) saves
//                       argmin, second argmin, the sign bits of L and the
//                       mask of slots whose clip was active; writes
//                       dLoss/dP of the final posteriors (and, with the
//                       per-iteration loss, of the posteriors after each
//                       earlier iteration) and per-workgroup loss sums
//   learned_bwd_kernel  reverse sweep (t = T-1..0, layers in reverse); the
//                       gradient of a slot's magnitude flows to the argmin
//                       (or, for the argmin slot, the second argmin) with the
//                       sign of that L; per-workgroup offset-gradient partials
//   reduce / adam       fixed-order reduction of the partials (deterministic)
//                       and an Adam step
// The forward's arithmetic equals MODE_LEARNED of tdmp_kernel bit for bit.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "learned.hpp"
#include "qc_code.hpp"

namespace wl {

constexpr int kTrainThreads = 512;
constexpr uint32_t kNoIdx = 31;

template <int RATE, int Z>
__global__ __launch_bounds__(kTrainThreads) void learned_fwd_kernel(TrainArgs a)
{
    using CD = Code<RATE, Z>;
    constexpr int NL = CD::NL, N = CD::N, C = CD::C;
    extern __shared__ __align__(16) unsigned char smem[];
    float* red = reinterpret_cast<float*>(smem);                 // [16] wave partials
    float* rxall = reinterpret_cast<float*>(smem + kTrainLdsPad);
    const int tid = threadIdx.x, F = a.frames_per_wg;
    const unsigned long long frame0 = (unsigned long long)blockIdx.x * F;
    const unsigned long long left = a.nframes - frame0;
    const int Feff = left < (unsigned long long)F ? (int)left : F;
    const int f = tid / Z, r = tid - f * Z;
    const bool active = f < Feff;
    for (int i = tid; i < Feff * N; i += blockDim.x) rxall[i] = a.llr[frame0 * N + i];
    __syncthreads();
    float* const base0 = rxall + (active ? f : 0) * N + r;
    float* const baseW = base0 - Z;
    auto at = [&](auto Lc, auto Jc) -> float* {
        constexpr int L = decltype(Lc)::value, J = decltype(Jc)::value;
        constexpr int b = CD::blk(L, J), s = CD::sh(L, J);
        if constexpr (s == 0) return base0 + b * Z;
        else return ((r >= Z - s) ? baseW : base0) + (b * Z + s);
    };
    float msg[NL][C];
    static_for<0, NL>([&](auto Lc) {
        static_for<0, C>([&](auto Jc) { msg[decltype(Lc)::value][decltype(Jc)::value] = 0.0f; });
    });
    double lsum_it = 0.0;   // this thread's part of the per-iteration loss terms
    for (int t = 0; t < a.T; t++) {
        static_for<0, NL>([&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            constexpr int D = CD::deg(L);
            if (active) {
                float x[D];
                float mn1 = 10000.0f, mn2 = 10000.0f;
                uint32_t i1 = kNoIdx, i2 = kNoIdx, sw = 0;
                static_for<0, D>([&](auto Jc) {
                    constexpr int J = decltype(Jc)::value;
                    const float v = *at(Lc, Jc) - msg[L][J];
                    x[J] = v;
                    const float av = fabsf(v);
                    const bool c1 = av < mn1, c2 = av < mn2;
                    mn2 = c1 ? mn1 : (c2 ? av : mn2);
                    i2 = c1 ? i1 : (c2 ? (uint32_t)J : i2);
                    mn1 = c1 ? av : mn1;
                    i1 = c1 ? (uint32_t)J : i1;
                    sw |= (__float_as_uint(v) >> 31) << J;
                });
                const uint32_t kw = ((uint32_t)(__popc(sw) & 1) ^ (uint32_t)(D & 1)) << 31;
                const uint32_t am = i1 == kNoIdx ? 0u : i1;
                const float* beta = a.offsets + ((size_t)t * NL + L) * C;
                uint32_t act = 0;
                static_for<0, D>([&](auto Jc) {
                    constexpr int J = decltype(Jc)::value;
                    const float v = x[J];
                    const uint32_t sg = (__float_as_uint(v) & 0x80000000u) ^ kw;
                    float q = ((uint32_t)J == am ? mn2 : mn1) - beta[J];
                    const bool on = q > 0.0f;
                    q = on ? q : 0.0f;
                    act |= (uint32_t)on << J;
                    const float m = __uint_as_float(__float_as_uint(q) | sg);
                    msg[L][J] = m;
                    *at(Lc, Jc) = v + m;
                });
                const size_t si = (((frame0 + f) * (size_t)a.T + t) * NL + L) * Z + r;
                a.state[si] = make_uint2(i1 | (i2 << 5) | (act << 10), sw);
            }
            __syncthreads();
        });
        if (a.gpost_it && t < a.T - 1) {
            // the weighted loss term of iteration t and its dLoss/dP
            for (int i = tid; i < Feff * N; i += blockDim.x) {
                const int fi = i / N, v = i - fi * N;
                const float p = rxall[i];
                const float b = a.bits[frame0 * N + i] ? 1.0f : 0.0f;
                const float sp = fmaxf(p, 0.0f) + log1pf(expf(-fabsf(p)));
                lsum_it += (double)(a.w_it * (sp - b * p));
                const float sg = 1.0f / (1.0f + expf(-p));
                a.gpost_it[((frame0 + fi) * (size_t)(a.T - 1) + t) * N + v] = (sg - b) * a.grad_scale * a.w_it;
            }
            __syncthreads();   // rxall is read here and rewritten by the next iteration
        }
    }
    // loss and dLoss/dP of the final posteriors
    double lsum = lsum_it;
    for (int i = tid; i < Feff * N; i += blockDim.x) {
        const float p = rxall[i];
        const float b = a.bits[frame0 * N + i] ? 1.0f : 0.0f;
        const float sp = fmaxf(p, 0.0f) + log1pf(expf(-fabsf(p)));
        lsum += (double)(a.w_final * (sp - b * p));
        const float sg = 1.0f / (1.0f + expf(-p));
        a.gpost[frame0 * N + i] = (sg - b) * a.grad_scale * a.w_final;
        if (a.post) a.post[frame0 * N + i] = p;
    }
    float ls = (float)lsum;
    for (int o = 32; o > 0; o >>= 1) ls += __shfl_xor(ls, o);
    if ((tid & 63) == 0) red[tid >> 6] = ls;
    __syncthreads();
    if (tid == 0) {
        float s = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
        a.loss_part[blockIdx.x] = s;
    }
}

template <int RATE, int Z>
__global__ __launch_bounds__(kTrainThreads) void learned_bwd_kernel(TrainArgs a)
{
    using CD = Code<RATE, Z>;
    constexpr int NL = CD::NL, N = CD::N, C = CD::C;
    extern __shared__ __align__(16) unsigned char smem[];
    float* red = reinterpret_cast<float*>(smem + kTrainLdsPad);  // [waves][NL*C]
    const int nw = blockDim.x >> 6;
    float* gall = red + nw * NL * C;                             // [F][N] dLoss/dP
    const int tid = threadIdx.x, F = a.frames_per_wg;
    const unsigned long long frame0 = (unsigned long long)blockIdx.x * F;
    const unsigned long long left = a.nframes - frame0;
    const int Feff = left < (unsigned long long)F ? (int)left : F;
    const int f = tid / Z, r = tid - f * Z;
    const bool active = f < Feff;
    for (int i = tid; i < Feff * N; i += blockDim.x) gall[i] = a.gpost[frame0 * N + i];
    __syncthreads();
    float* const base0 = gall + (active ? f : 0) * N + r;
    float* const baseW = base0 - Z;
    auto at = [&](auto Lc, auto Jc) -> float* {
        constexpr int L = decltype(Lc)::value, J = decltype(Jc)::value;
        constexpr int b = CD::blk(L, J), s = CD::sh(L, J);
        if constexpr (s == 0) return base0 + b * Z;
        else return ((r >= Z - s) ? baseW : base0) + (b * Z + s);
    };
    float gm[NL][C];
    static_for<0, NL>([&](auto Lc) {
        static_for<0, C>([&](auto Jc) { gm[decltype(Lc)::value][decltype(Jc)::value] = 0.0f; });
    });
    for (int t = a.T - 1; t >= 0; t--) {
        if (a.gpost_it && t < a.T - 1) {
            // the loss term of iteration t joins the gradient of its posteriors
            for (int i = tid; i < Feff * N; i += blockDim.x) {
                const int fi = i / N, v = i - fi * N;
                gall[i] += a.gpost_it[((frame0 + fi) * (size_t)(a.T - 1) + t) * N + v];
            }
            __syncthreads();
        }
        float db[NL][C];
        static_for<0, NL>([&](auto Lc) {
            static_for<0, C>([&](auto Jc) { db[decltype(Lc)::value][decltype(Jc)::value] = 0.0f; });
        });
        static_for<0, NL>([&](auto Lrev) {
            constexpr int L = NL - 1 - decltype(Lrev)::value;
            constexpr int D = CD::deg(L);
            if (active) {
                const uint2 st = a.state[(((frame0 + f) * (size_t)a.T + t) * NL + L) * Z + r];
                const uint32_t i1 = st.x & 31u, i2 = (st.x >> 5) & 31u, act = st.x >> 10, sw = st.y;
                const uint32_t kbit = ((uint32_t)(__popc(sw) & 1) ^ (uint32_t)(D & 1));
                const uint32_t am = i1 == kNoIdx ? 0u : i1;
                float dl[D], dm[D];
                float s1 = 0.0f, s2 = 0.0f;
                static_for<0, D>([&](auto Jc) {
                    constexpr int J = decltype(Jc)::value;
                    const float g = *at(std::integral_constant<int, L>{}, Jc);
                    const float dmsg = g + gm[L][J];
                    dl[J] = g;
                    const bool neg = (((sw >> J) & 1u) ^ kbit) != 0;   // message sign
                    const float dmag = ((act >> J) & 1u) ? (neg ? -dmsg : dmsg) : 0.0f;
                    dm[J] = dmag;
                    db[L][J] -= dmag;
                    if ((uint32_t)J == am) s2 += dmag;
                    else s1 += dmag;
                });
                static_for<0, D>([&](auto Jc) {
                    constexpr int J = decltype(Jc)::value;
                    const float sgn = ((sw >> J) & 1u) ? -1.0f : 1.0f;   // d|L|/dL
                    float d = dl[J];
                    if ((uint32_t)J == i1) d += sgn * s1;
                    if ((uint32_t)J == i2) d += sgn * s2;
                    *at(std::integral_constant<int, L>{}, Jc) = d;
                    gm[L][J] = -d;
                });
            }
            __syncthreads();
        });
        // deterministic reduction of this iteration's offset gradients
        static_for<0, NL>([&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            static_for<0, C>([&](auto Jc) {
                constexpr int J = decltype(Jc)::value;
                float v = db[L][J];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                if ((tid & 63) == 0) red[(tid >> 6) * NL * C + L * C + J] = v;
            });
        });
        __syncthreads();
        for (int p = tid; p < NL * C; p += blockDim.x) {
            float s = 0.0f;
            for (int w = 0; w < nw; w++) s += red[w * NL * C + p];
            a.dbeta_part[((size_t)blockIdx.x * a.T + t) * NL * C + p] = s;
        }
        __syncthreads();
    }
}

// sum of the per-workgroup partials, one thread per parameter, fixed order
__global__ void reduce_partials_kernel(const float* __restrict__ part, int nparts, int nparam,
                                       float* __restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nparam) return;
    float s = 0.0f;
    for (int w = 0; w < nparts; w++) s += part[(size_t)w * nparam + p];
    out[p] = s;
}

// Adam (Kingma & Ba 2015) with bias correction, one workgroup.  The 1-based
// step is *step_dev + 1; the kernel then advances *step_dev, so a captured
// training step replays with the right count.
__global__ __launch_bounds__(1024) void adam_kernel(float* __restrict__ theta, const float* __restrict__ g,
                                                    float* __restrict__ m, float* __restrict__ v, int nparam,
                                                    float lr, float b1, float b2, float eps,
                                                    unsigned long long* __restrict__ step_dev)
{
    const unsigned long long step = *step_dev + 1;
    const float c1 = 1.0f - powf(b1, (float)step);
    const float c2 = 1.0f - powf(b2, (float)step);
    for (int p = threadIdx.x; p < nparam; p += blockDim.x) {
        const float gi = g[p];
        const float mi = b1 * m[p] + (1.0f - b1) * gi;
        const float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
        m[p] = mi;
        v[p] = vi;
        theta[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
    __syncthreads();
    if (threadIdx.x == 0) *step_dev = step;
}

template <int RATE, int Z>
hipError_t launch_learned_fwd(const TrainArgs& a, hipStream_t s)
{
    const int F = a.frames_per_wg;
    const unsigned long long grid = (a.nframes + F - 1) / F;
    const int threads = ((F * Z + 63) / 64) * 64;
    const size_t lds = kTrainLdsPad + (size_t)F * Code<RATE, Z>::N * sizeof(float);
    hipLaunchKernelGGL((learned_fwd_kernel<RATE, Z>), dim3((unsigned)grid), dim3(threads), lds, s, a);
    return hipGetLastError();
}

template <int RATE, int Z>
hipError_t launch_learned_bwd(const TrainArgs& a, hipStream_t s)
{
    using CD = Code<RATE, Z>;
    const int F = a.frames_per_wg;
    const unsigned long long grid = (a.nframes + F - 1) / F;
    const int threads = ((F * Z + 63) / 64) * 64;
    const size_t lds = kTrainLdsPad + (size_t)(threads / 64) * CD::NL * CD::C * sizeof(float) +
                       (size_t)F * CD::N * sizeof(float);
    hipLaunchKernelGGL((learned_bwd_kernel<RATE, Z>), dim3((unsigned)grid), dim3(threads), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const float* part, int nparts, int nparam, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((nparam + 255) / 256), dim3(256), 0, s, part, nparts,
                       nparam, out);
    return hipGetLastError();
}

hipError_t launch_adam(float* theta, const float* g, float* m, float* v, int nparam, float lr, float b1,
                       float b2, float eps, unsigned long long* step_dev, hipStream_t s)
{
    hipLaunchKernelGGL(adam_kernel, dim3(1), dim3(1024), 0, s, theta, g, m, v, nparam, lr, b1, b2, eps, step_dev);
    return hipGetLastError();
}

template <int RATE, int Z>
constexpr int train_frames_per_wg()
{
    // one workgroup of up to 192 threads, whole frames
    int f = 192 / Z;
    return f < 1 ? 1 : f;
}

#define WL_TRAIN_ENTRY(R, Zv) \
    { R, Zv, &launch_learned_fwd<R, Zv>, &launch_learned_bwd<R, Zv>, train_frames_per_wg<R, Zv>() }

const TrainEntry kTrainTable[] = {
    WL_TRAIN_ENTRY(0, 24), WL_TRAIN_ENTRY(1, 24), WL_TRAIN_ENTRY(2, 24), WL_TRAIN_ENTRY(3, 24),
    WL_TRAIN_ENTRY(4, 24), WL_TRAIN_ENTRY(5, 24),
    WL_TRAIN_ENTRY(0, 96), WL_TRAIN_ENTRY(1, 96), WL_TRAIN_ENTRY(2, 96), WL_TRAIN_ENTRY(3, 96),
    WL_TRAIN_ENTRY(4, 96), WL_TRAIN_ENTRY(5, 96),
};

const TrainEntry* train_entry(int rate, int z)
{
    for (const TrainEntry& e : kTrainTable)
        if (e.rate == rate && e.z == z) return &e;
    return nullptr;
}

}  // namespace wl
