// generic.hip -- the layered normalised min-sum decoder for a quasi-cyclic
// code given at run time (generic.hpp).
//
// Same arithmetic as the compiled decoder (tdmp.hpp), which restates
// lib/ldpc_decoder.cc:122-262 bit-exactly; here the code shape is read from a
// device descriptor instead of being compiled in, so slot loops run over
// run-time degrees and the check-to-variable messages live in LDS next to the
// frame's posteriors (messages of layer l, slot j, row r at moff[l] + j z + r,
// consecutive rows in consecutive banks).  One frame per workgroup, lane r
// owns row r of every layer.  Between the two passes of a row the extrinsic
// values L = rx - msg are parked in the message slots.
#include <hip/hip_runtime.h>

#include "generic.hpp"

namespace wl {

__global__ __launch_bounds__(kGenMaxZ) void generic_kernel(GenArgs a)
{
    extern __shared__ __align__(16) unsigned char gsm[];
    const GenDesc& d = *a.desc;
    const int z = d.z, n = d.n, mb = d.mb;
    const unsigned long long frame = blockIdx.x;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(gsm);
    float* rx = reinterpret_cast<float*>(gsm + kGenLdsPad);
    float* msg = rx + n;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int r = tid;
    const bool act = r < z;

    const float* src = a.llr + frame * (unsigned long long)n;
    for (int i = tid; i < n; i += nt) rx[i] = src[i];
    for (int i = tid; i < d.edges; i += nt) msg[i] = 0.0f;
    if (tid == 0) cnt[0] = 0;
    __syncthreads();

    // variable of slot (l, j) for row r
    auto var = [&](int l, int j) -> int {
        int c = r + d.sh[l][j];
        if (c >= z) c -= z;
        return d.blk[l][j] * z + c;
    };

    // failing parity checks among this lane's rows over hd = rx >= 0
    // (lib/ldpc_decoder.cc:67-104)
    auto fail_rows = [&]() -> int {
        if (!act) return 0;
        int fails = 0;
        for (int l = 0; l < mb; l++) {
            uint32_t p = 0;
            const int dg = d.deg[l];
            for (int j = 0; j < dg; j++) p ^= rx[var(l, j)] >= 0.0f ? 1u : 0u;
            fails += (int)p;
        }
        return fails;
    };

    // whether some row of the frame fails, for the stop decisions (compute_
    // syndrome(hd, true) returns at the first failing row): a lane sets
    // cnt[0] at its first failing row and stops at it, or as soon as it sees
    // the flag set (by itself or by another lane; the decision is read after
    // a barrier)
    auto mark_fail = [&]() {
        if (!act) return;
        const volatile uint32_t* flag = cnt;   // re-read every layer
        for (int l = 0; l < mb && *flag == 0; l++) {
            uint32_t p = 0;
            const int dg = d.deg[l];
            for (int j = 0; j < dg; j++) p ^= rx[var(l, j)] >= 0.0f ? 1u : 0u;
            if (p) {
                cnt[0] = 1;
                break;
            }
        }
    };

    // one layer for this lane's row (lib/ldpc_decoder.cc:143-240)
    auto layer = [&](int l) {
        const int dg = d.deg[l];
        float* ml = msg + d.moff[l] + r;
        float mn1 = 10000.0f, mn2 = 10000.0f;
        uint32_t sx = 0;
        for (int j = 0; j < dg; j++) {
            const float v = rx[var(l, j)] - ml[j * z];
            ml[j * z] = v;
            const float av = fabsf(v);
            mn2 = __builtin_amdgcn_fmed3f(mn1, mn2, av);
            mn1 = fminf(mn1, av);
            sx ^= __float_as_uint(v);
        }
        // message sign bit of slot n = signbit(L_n) ^ parity(signbits) ^ (deg & 1)
        const uint32_t kw = (sx ^ ((uint32_t)(dg & 1) << 31)) & 0x80000000u;
        const double d1 = __dmul_rn((double)mn1, 0.815);
        const double d2 = __dmul_rn((double)mn2, 0.815);
        const uint32_t h1 = (uint32_t)__double2hiint(d1) | kw, h2 = (uint32_t)__double2hiint(d2) | kw;
        const int l1 = __double2loint(d1), l2 = __double2loint(d2);
        for (int j = 0; j < dg; j++) {
            const float v = ml[j * z];
            const bool e = fabsf(v) == mn1;
            const uint32_t hi = (e ? h2 : h1) ^ (__float_as_uint(v) & 0x80000000u);
            const double dsg = __hiloint2double((int)hi, e ? l2 : l1);
            ml[j * z] = __double2float_rn(dsg);
            rx[var(l, j)] = __double2float_rn(__dadd_rn((double)v, dsg));
        }
    };

    bool done = false;
    int it_done = 0;
    if (a.early_stop) {
        // initial hard decision + syndrome (lib/ldpc_decoder.cc:130-139)
        mark_fail();
        __syncthreads();
        done = cnt[0] == 0;
        __syncthreads();
        if (tid == 0) cnt[0] = 0;
        __syncthreads();
    }
    for (int it = 0; it < a.max_iter && !done; it++) {   // done is uniform: one frame
        for (int l = 0; l < mb; l++) {
            if (act) layer(l);
            __syncthreads();
        }
        if (a.early_stop) {
            // lib/ldpc_decoder.cc:242-251
            mark_fail();
            __syncthreads();
            if (cnt[0] == 0) {
                done = true;
                it_done = it + 1;
            }
            __syncthreads();
            if (tid == 0) cnt[0] = 0;
            __syncthreads();
        }
    }
    const bool converged = a.early_stop && done;
    if (!converged) it_done = a.max_iter;

    // Final syndrome: the weight where it is asked for (a frame that stopped
    // early satisfies every check: weight 0); with fixed iterations the
    // return value only needs to know whether a row fails; with early stop
    // it follows from the stop decisions.
    if (a.synd) {
        const int fr = converged ? 0 : fail_rows();
        if (fr) atomicAdd(&cnt[0], (uint32_t)fr);
    } else if (a.ret && !a.early_stop) {
        mark_fail();
    }
    __syncthreads();
    if (tid == 0) {
        const uint32_t w = cnt[0];
        if (a.synd) a.synd[frame] = (uint16_t)w;
        if (a.iters) a.iters[frame] = (uint8_t)it_done;
        if (a.ret) {
            if (a.early_stop) a.ret[frame] = (converged || a.max_iter == 0) ? 0 : 1;   // :124,141,261
            else a.ret[frame] = w != 0;
        }
    }
    if (a.post) {
        float* dst = a.post + frame * (unsigned long long)n;
        for (int i = tid; i < n; i += nt) dst[i] = rx[i];
    }
    if (a.hard) {
        // ceil(n / 8) bytes per frame, bit v % 8 of byte v / 8, padding bits 0
        const int nb8 = (n + 7) / 8;
        uint8_t* dst = a.hard + frame * (unsigned long long)nb8;
        for (int q = tid; q < nb8; q += nt) {
            uint32_t byte = 0;
            for (int b = 0; b < 8; b++) {
                const int v = q * 8 + b;
                if (v < n && rx[v] >= 0.0f) byte |= 1u << b;
            }
            dst[q] = (uint8_t)byte;
        }
    }
}

hipError_t launch_generic(const GenArgs& a, hipStream_t stream)
{
    if (a.nframes == 0) return hipSuccess;
    if (a.nframes > 0x7fffffffull || a.z < 1 || a.z > kGenMaxZ) return hipErrorInvalidValue;
    const size_t lds = gen_lds_bytes(a.n, a.edges);
    if (lds > kGenMaxLds) return hipErrorInvalidValue;
    const int threads = ((a.z + 63) / 64) * 64;
    hipLaunchKernelGGL(generic_kernel, dim3((unsigned)a.nframes), dim3(threads), lds, stream, a);
    return hipGetLastError();
}

}  // namespace wl
