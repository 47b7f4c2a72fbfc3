// learned.hpp -- training-kernel interface of the learned offset min-sum decoder.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wl {

constexpr int kTrainLdsPad = 256;

struct TrainArgs {
    const float* llr;        // [nframes][n] reference-convention LLRs
    const uint8_t* bits;     // [nframes][n] transmitted codeword bits (targets)
    const float* offsets;    // [T][NL][C]
    uint2* state;            // [nframes][T][NL][z] saved selection state
    float* gpost;            // [nframes][n] dLoss/dP of the final posteriors
    float* post;             // [nframes][n] final posteriors (nullable)
    float* loss_part;        // [workgroups] loss sums
    float* dbeta_part;       // [workgroups][T][NL][C] offset-gradient partials
    unsigned long long nframes;
    int frames_per_wg;
    int T;
    float grad_scale;        // 1 / (nframes * n): the loss is a mean
    // per-iteration loss: the weighted sum of the BCE after every unrolled
    // iteration; w_final weights the last one, w_it each of the T-1 others
    // (w_it = 0, gpost_it unused: the final posteriors only)
    float* gpost_it;         // [nframes][T-1][n] dLoss/dP after iterations 0 .. T-2 (nullable)
    float w_final, w_it;
};

using TrainLaunch = hipError_t (*)(const TrainArgs&, hipStream_t);

struct TrainEntry {
    int rate, z;
    TrainLaunch fwd, bwd;
    int frames_per_wg;
};

const TrainEntry* train_entry(int rate, int z);
hipError_t launch_reduce_partials(const float* part, int nparts, int nparam, float* out, hipStream_t s);
hipError_t launch_adam(float* theta, const float* g, float* m, float* v, int nparam, float lr, float b1,
                       float b2, float eps, unsigned long long* step_dev, hipStream_t s);

}  // namespace wl
