// trainer.hip -- C ABI of the learned offset min-sum trainer.
//
// One training step on the GPU:
//   channel_kernel  fresh seeded frames (Philox counter = step * batch + i),
//                   random Eb/N0 per frame, encoder, BPSK/AWGN, LLRs + bits
//   learned_fwd     T unrolled iterations, saved selection state, dLoss/dP
//   learned_bwd     reverse sweep, per-workgroup offset-gradient partials
//   reduce          fixed-order sum of the partials (deterministic)
//   adam            Adam step; advances the device step counter
// The step's launches are captured once in a hipGraph and replayed, so a
// training run is one graph launch per step with no host work in between.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/wimax_ldpc_amd.h"
#include "encoder.hpp"
#include "errors.hpp"
#include "learned.hpp"
#include "qc_code.hpp"

namespace {

// The C ABI's error text lives in capi.hip; this file reports through it.
int tfail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    wl::set_error_text(buf);
    return code;
}

#define TR_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return tfail(WL_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(T** p, size_t n)
{
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) return tfail(WL_ENOMEM, "hipMalloc(%zu): %s", n * sizeof(T), hipGetErrorString(e));
    return WL_OK;
}

const uint32_t kCkptMagic = 0x52544c57u;  // "WLTR"
const uint32_t kCkptVersion = 1;

}  // namespace

struct wl_trainer {
    int rate = 0, z = 0, T = 0, batch = 0, nl = 0, c = 0, n = 0, m = 0, k = 0, nparam = 0;
    uint64_t seed = 0;
    float lr = 1e-3f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float eb_lo = 1.0f, eb_hi = 4.0f;
    const wl::TrainEntry* entry = nullptr;
    wl::EncodeDesc enc{};
    int nwg = 0;
    float *theta = nullptr, *mom = nullptr, *vel = nullptr, *grad = nullptr;
    float *llr = nullptr, *gpost = nullptr, *loss_part = nullptr, *dbeta_part = nullptr, *loss_sum = nullptr;
    uint8_t* bits = nullptr;
    uint2* state = nullptr;
    unsigned long long* step_dev = nullptr;
    // scratch of wl_trainer_grad_dev, grown on demand
    size_t gcap = 0;
    uint2* gstate = nullptr;
    float *ggpost = nullptr, *gloss = nullptr, *gdb = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    float g_lo = 0.0f, g_hi = 0.0f;   // Eb/N0 range the graph was captured with
    // per-iteration loss: dLoss/dP after iterations 0 .. T-2, training batch
    // and wl_trainer_grad_dev scratch
    int per_iter = 0;
    float* gpost_it = nullptr;
    float* ggpost_it = nullptr;
    size_t gcap_it = 0;
};

namespace {

void free_all(wl_trainer* t)
{
    if (t->gexec) (void)hipGraphExecDestroy(t->gexec);
    if (t->graph) (void)hipGraphDestroy(t->graph);
    void* ps[] = {t->theta, t->mom, t->vel, t->grad, t->llr, t->gpost, t->loss_part, t->dbeta_part,
                  t->loss_sum, t->bits, t->state, t->step_dev, t->gstate, t->ggpost, t->gloss, t->gdb,
                  t->gpost_it, t->ggpost_it};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (t->stream) (void)hipStreamDestroy(t->stream);
}

// loss weights: the final posteriors only, or 1/T for each iteration's
void set_loss_weights(const wl_trainer* t, wl::TrainArgs& a, float* gpost_it)
{
    a.gpost_it = (t->per_iter && t->T > 1) ? gpost_it : nullptr;
    a.w_final = a.gpost_it ? 1.0f / (float)t->T : 1.0f;
    a.w_it = a.gpost_it ? 1.0f / (float)t->T : 0.0f;
}

int enqueue_step(wl_trainer* t, hipStream_t s)
{
    hipError_t e = wl::launch_channel(t->enc, t->seed, 0, (unsigned long long)t->batch, t->eb_lo, t->eb_hi,
                                      t->llr, t->bits, nullptr, s, t->step_dev);
    if (e != hipSuccess) return tfail(WL_EDEVICE, "channel launch: %s", hipGetErrorString(e));
    wl::TrainArgs a{};
    a.llr = t->llr;
    a.bits = t->bits;
    a.offsets = t->theta;
    a.state = t->state;
    a.gpost = t->gpost;
    a.post = nullptr;
    a.loss_part = t->loss_part;
    a.dbeta_part = t->dbeta_part;
    a.nframes = (unsigned long long)t->batch;
    a.frames_per_wg = t->entry->frames_per_wg;
    a.T = t->T;
    a.grad_scale = 1.0f / ((float)t->batch * (float)t->n);
    set_loss_weights(t, a, t->gpost_it);
    if ((e = t->entry->fwd(a, s)) != hipSuccess) return tfail(WL_EDEVICE, "forward: %s", hipGetErrorString(e));
    if ((e = t->entry->bwd(a, s)) != hipSuccess) return tfail(WL_EDEVICE, "backward: %s", hipGetErrorString(e));
    if ((e = wl::launch_reduce_partials(t->dbeta_part, t->nwg, t->nparam, t->grad, s)) != hipSuccess)
        return tfail(WL_EDEVICE, "reduce: %s", hipGetErrorString(e));
    if ((e = wl::launch_reduce_partials(t->loss_part, t->nwg, 1, t->loss_sum, s)) != hipSuccess)
        return tfail(WL_EDEVICE, "reduce: %s", hipGetErrorString(e));
    if ((e = wl::launch_adam(t->theta, t->grad, t->mom, t->vel, t->nparam, t->lr, t->b1, t->b2, t->eps,
                             t->step_dev, s)) != hipSuccess)
        return tfail(WL_EDEVICE, "adam: %s", hipGetErrorString(e));
    return WL_OK;
}

}  // namespace

extern "C" {

int wl_trainer_create(int rate, int z, int T, int batch, uint64_t seed, float lr, float init_offset,
                      wl_trainer** out)
{
    if (!out) return tfail(WL_EINVAL, "null output");
    *out = nullptr;
    if (rate < 0 || rate > 5 || !wl::valid_z(z)) return tfail(WL_EINVAL, "invalid rate/z %d/%d", rate, z);
    if (rate == 4) return tfail(WL_ENOTSUP, "3/4B has no valid encoder (training frames need one)");
    const wl::TrainEntry* en = wl::train_entry(rate, z);
    if (!en) return tfail(WL_ENOTSUP, "training kernels are built for z = 24 and z = 96 only (z = %d)", z);
    if (T < 1 || T > 32) return tfail(WL_EINVAL, "T %d out of range 1..32", T);
    if (batch < 1 || batch > (1 << 22)) return tfail(WL_EINVAL, "batch %d out of range", batch);
    wl_trainer* t = new (std::nothrow) wl_trainer();
    if (!t) return tfail(WL_ENOMEM, "out of host memory");
    const wl::LayerTable lt = wl::make_layer_table(rate);
    const wl_base_info& bi = wl_base_table[rate];
    t->rate = rate;
    t->z = z;
    t->T = T;
    t->batch = batch;
    t->nl = lt.nl;
    t->c = lt.c;
    t->n = 24 * z;
    t->m = lt.nl * z;
    t->k = t->n - t->m;
    t->nparam = T * lt.nl * lt.c;
    t->seed = seed;
    t->lr = lr;
    t->entry = en;
    t->nwg = (batch + en->frames_per_wg - 1) / en->frames_per_wg;
    wl::EncodeDesc& d = t->enc;
    d.z = z;
    d.n = t->n;
    d.k = t->k;
    d.m = t->m;
    d.mb = lt.nl;
    d.kb = 24 - lt.nl;
    d.x = bi.x;
    d.d = wl::lift(rate, bi.d0, z);
    d.frames_per_wg = 256 / z > 0 ? 256 / z : 1;
    for (int l = 0; l < lt.nl; l++) {
        d.deg[l] = (int8_t)lt.deg[l];
        for (int j = 0; j < lt.deg[l]; j++) {
            d.blk[l][j] = (int8_t)lt.blk[l][j];
            d.sh[l][j] = (int8_t)wl::lift(rate, lt.p96[l][j], z);
        }
    }
    int rc;
    if ((rc = dalloc(&t->theta, t->nparam)) || (rc = dalloc(&t->mom, t->nparam)) ||
        (rc = dalloc(&t->vel, t->nparam)) || (rc = dalloc(&t->grad, t->nparam)) ||
        (rc = dalloc(&t->llr, (size_t)batch * t->n)) || (rc = dalloc(&t->gpost, (size_t)batch * t->n)) ||
        (rc = dalloc(&t->bits, (size_t)batch * t->n)) || (rc = dalloc(&t->loss_part, t->nwg)) ||
        (rc = dalloc(&t->dbeta_part, (size_t)t->nwg * t->nparam)) || (rc = dalloc(&t->loss_sum, 1)) ||
        (rc = dalloc(&t->state, (size_t)batch * T * t->m)) || (rc = dalloc(&t->step_dev, 1))) {
        free_all(t);
        delete t;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free_all(t);
        delete t;
        return tfail(WL_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    std::vector<float> init(t->nparam, init_offset), zero(t->nparam, 0.0f);
    unsigned long long z0 = 0;
    if (hipMemcpy(t->theta, init.data(), t->nparam * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->mom, zero.data(), t->nparam * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->vel, zero.data(), t->nparam * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->step_dev, &z0, sizeof z0, hipMemcpyHostToDevice) != hipSuccess) {
        free_all(t);
        delete t;
        return tfail(WL_EDEVICE, "trainer initialisation copy failed");
    }
    *out = t;
    return WL_OK;
}

int wl_trainer_destroy(wl_trainer* t)
{
    if (!t) return WL_OK;
    free_all(t);
    delete t;
    return WL_OK;
}

int wl_trainer_set_hparams(wl_trainer* t, float lr, float beta1, float beta2, float eps)
{
    if (!t) return tfail(WL_EINVAL, "null trainer");
    if (!(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return tfail(WL_EINVAL, "invalid Adam hyper-parameters");
    t->lr = lr;
    t->b1 = beta1;
    t->b2 = beta2;
    t->eps = eps;
    if (t->gexec) {   // captured launches hold the old values
        (void)hipGraphExecDestroy(t->gexec);
        (void)hipGraphDestroy(t->graph);
        t->gexec = nullptr;
        t->graph = nullptr;
    }
    return WL_OK;
}

int wl_trainer_set_loss(wl_trainer* t, int per_iteration)
{
    if (!t) return tfail(WL_EINVAL, "null trainer");
    if (per_iteration != 0 && per_iteration != 1) return tfail(WL_EINVAL, "per_iteration is 0 or 1");
    if (per_iteration && t->T > 1 && !t->gpost_it) {
        int rc = dalloc(&t->gpost_it, (size_t)t->batch * (t->T - 1) * t->n);
        if (rc) return rc;
    }
    t->per_iter = per_iteration;
    if (t->gexec) {   // the captured step holds the old loss
        (void)hipGraphExecDestroy(t->gexec);
        (void)hipGraphDestroy(t->graph);
        t->gexec = nullptr;
        t->graph = nullptr;
    }
    return WL_OK;
}

int wl_trainer_step(wl_trainer* t, int nsteps, float ebn0_lo, float ebn0_hi, float* last_loss)
{
    if (!t) return tfail(WL_EINVAL, "null trainer");
    if (nsteps < 0) return tfail(WL_EINVAL, "nsteps < 0");
    if (t->gexec && (ebn0_lo != t->g_lo || ebn0_hi != t->g_hi)) {
        (void)hipGraphExecDestroy(t->gexec);
        (void)hipGraphDestroy(t->graph);
        t->gexec = nullptr;
        t->graph = nullptr;
    }
    t->eb_lo = ebn0_lo;
    t->eb_hi = ebn0_hi;
    if (!t->gexec && nsteps > 0) {
        TR_HIP(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step(t, t->stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(t->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return tfail(WL_EDEVICE, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&t->gexec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return tfail(WL_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
        }
        t->graph = g;
        t->g_lo = ebn0_lo;
        t->g_hi = ebn0_hi;
    }
    for (int i = 0; i < nsteps; i++) TR_HIP(hipGraphLaunch(t->gexec, t->stream));
    TR_HIP(hipStreamSynchronize(t->stream));
    if (last_loss) {
        float s = 0.0f;
        TR_HIP(hipMemcpy(&s, t->loss_sum, sizeof s, hipMemcpyDeviceToHost));
        *last_loss = s / ((float)t->batch * (float)t->n);
    }
    return WL_OK;
}

int wl_trainer_num_params(const wl_trainer* t, int* T, int* mb, int* c)
{
    if (!t || !T || !mb || !c) return tfail(WL_EINVAL, "null argument");
    *T = t->T;
    *mb = t->nl;
    *c = t->c;
    return WL_OK;
}

int wl_trainer_get_offsets(wl_trainer* t, float* out)
{
    if (!t || !out) return tfail(WL_EINVAL, "null argument");
    TR_HIP(hipStreamSynchronize(t->stream));
    TR_HIP(hipMemcpy(out, t->theta, t->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return WL_OK;
}

int wl_trainer_set_offsets(wl_trainer* t, const float* in)
{
    if (!t || !in) return tfail(WL_EINVAL, "null argument");
    TR_HIP(hipStreamSynchronize(t->stream));
    TR_HIP(hipMemcpy(t->theta, in, t->nparam * sizeof(float), hipMemcpyHostToDevice));
    return WL_OK;
}

int wl_trainer_save(wl_trainer* t, const char* path)
{
    if (!t || !path) return tfail(WL_EINVAL, "null argument");
    TR_HIP(hipStreamSynchronize(t->stream));
    std::vector<float> th(t->nparam), mo(t->nparam), ve(t->nparam);
    unsigned long long step = 0;
    TR_HIP(hipMemcpy(th.data(), t->theta, t->nparam * 4, hipMemcpyDeviceToHost));
    TR_HIP(hipMemcpy(mo.data(), t->mom, t->nparam * 4, hipMemcpyDeviceToHost));
    TR_HIP(hipMemcpy(ve.data(), t->vel, t->nparam * 4, hipMemcpyDeviceToHost));
    TR_HIP(hipMemcpy(&step, t->step_dev, sizeof step, hipMemcpyDeviceToHost));
    FILE* fp = fopen(path, "wb");
    if (!fp) return tfail(WL_EINVAL, "cannot open %s for writing", path);
    const uint32_t hdr[8] = {kCkptMagic, kCkptVersion, (uint32_t)t->rate, (uint32_t)t->z,
                             (uint32_t)t->T, (uint32_t)t->nl, (uint32_t)t->c, (uint32_t)t->batch};
    bool ok = fwrite(hdr, sizeof hdr, 1, fp) == 1 && fwrite(&t->seed, 8, 1, fp) == 1 &&
              fwrite(&step, 8, 1, fp) == 1 && fwrite(th.data(), 4, t->nparam, fp) == (size_t)t->nparam &&
              fwrite(mo.data(), 4, t->nparam, fp) == (size_t)t->nparam &&
              fwrite(ve.data(), 4, t->nparam, fp) == (size_t)t->nparam;
    ok = (fclose(fp) == 0) && ok;
    return ok ? WL_OK : tfail(WL_EINVAL, "short write to %s", path);
}

int wl_trainer_load(wl_trainer* t, const char* path)
{
    if (!t || !path) return tfail(WL_EINVAL, "null argument");
    FILE* fp = fopen(path, "rb");
    if (!fp) return tfail(WL_EINVAL, "cannot open %s", path);
    uint32_t hdr[8];
    uint64_t seed = 0;
    unsigned long long step = 0;
    std::vector<float> th(t->nparam), mo(t->nparam), ve(t->nparam);
    bool ok = fread(hdr, sizeof hdr, 1, fp) == 1 && hdr[0] == kCkptMagic && hdr[1] == kCkptVersion &&
              hdr[2] == (uint32_t)t->rate && hdr[3] == (uint32_t)t->z && hdr[4] == (uint32_t)t->T &&
              hdr[5] == (uint32_t)t->nl && hdr[6] == (uint32_t)t->c && fread(&seed, 8, 1, fp) == 1 &&
              fread(&step, 8, 1, fp) == 1 && fread(th.data(), 4, t->nparam, fp) == (size_t)t->nparam &&
              fread(mo.data(), 4, t->nparam, fp) == (size_t)t->nparam &&
              fread(ve.data(), 4, t->nparam, fp) == (size_t)t->nparam;
    fclose(fp);
    if (!ok) return tfail(WL_EINVAL, "%s is not a checkpoint of this trainer's code and T", path);
    // training frames are drawn at frame offset step * batch: another batch
    // size would continue a different frame stream
    if (hdr[7] != (uint32_t)t->batch)
        return tfail(WL_EINVAL, "%s was written with batch %u, this trainer has batch %d", path, hdr[7], t->batch);
    TR_HIP(hipStreamSynchronize(t->stream));
    TR_HIP(hipMemcpy(t->theta, th.data(), t->nparam * 4, hipMemcpyHostToDevice));
    TR_HIP(hipMemcpy(t->mom, mo.data(), t->nparam * 4, hipMemcpyHostToDevice));
    TR_HIP(hipMemcpy(t->vel, ve.data(), t->nparam * 4, hipMemcpyHostToDevice));
    TR_HIP(hipMemcpy(t->step_dev, &step, sizeof step, hipMemcpyHostToDevice));
    t->seed = seed;
    if (t->gexec) {   // the seed is a captured kernel argument
        (void)hipGraphExecDestroy(t->gexec);
        (void)hipGraphDestroy(t->graph);
        t->gexec = nullptr;
        t->graph = nullptr;
    }
    return WL_OK;
}

int wl_trainer_get_step(wl_trainer* t, uint64_t* step)
{
    if (!t || !step) return tfail(WL_EINVAL, "null argument");
    unsigned long long s = 0;
    TR_HIP(hipStreamSynchronize(t->stream));
    TR_HIP(hipMemcpy(&s, t->step_dev, sizeof s, hipMemcpyDeviceToHost));
    *step = s;
    return WL_OK;
}

int wl_trainer_grad_dev(wl_trainer* t, const float* llr, const uint8_t* bits, size_t nframes, float* post,
                        float* grad_host, double* loss_host)
{
    if (!t || !llr || !bits) return tfail(WL_EINVAL, "null argument");
    if (nframes == 0 || nframes > (1u << 24)) return tfail(WL_EINVAL, "nframes %zu out of range", nframes);
    const int F = t->entry->frames_per_wg;
    const size_t nwg = (nframes + F - 1) / F;
    if (nframes > t->gcap) {
        void* ps[] = {t->gstate, t->ggpost, t->gloss, t->gdb};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        t->gstate = nullptr;
        t->ggpost = t->gloss = t->gdb = nullptr;
        t->gcap = 0;
        int rc;
        if ((rc = dalloc(&t->gstate, nframes * t->T * t->m)) || (rc = dalloc(&t->ggpost, nframes * t->n)) ||
            (rc = dalloc(&t->gloss, nwg + 1)) || (rc = dalloc(&t->gdb, (nwg + 1) * t->nparam)))
            return rc;
        t->gcap = nframes;
    }
    wl::TrainArgs a{};
    a.llr = llr;
    a.bits = bits;
    a.offsets = t->theta;
    a.state = t->gstate;
    a.gpost = t->ggpost;
    a.post = post;
    a.loss_part = t->gloss;
    a.dbeta_part = t->gdb;
    a.nframes = nframes;
    a.frames_per_wg = F;
    a.T = t->T;
    a.grad_scale = (float)(1.0 / ((double)nframes * t->n));
    if (t->per_iter && t->T > 1 && nframes > t->gcap_it) {
        if (t->ggpost_it) (void)hipFree(t->ggpost_it);
        t->ggpost_it = nullptr;
        t->gcap_it = 0;
        int rc = dalloc(&t->ggpost_it, nframes * (t->T - 1) * t->n);
        if (rc) return rc;
        t->gcap_it = nframes;
    }
    set_loss_weights(t, a, t->ggpost_it);
    hipError_t e;
    if ((e = t->entry->fwd(a, t->stream)) != hipSuccess) return tfail(WL_EDEVICE, "forward: %s", hipGetErrorString(e));
    if ((e = t->entry->bwd(a, t->stream)) != hipSuccess) return tfail(WL_EDEVICE, "backward: %s", hipGetErrorString(e));
    TR_HIP(hipStreamSynchronize(t->stream));
    std::vector<float> part(nwg * t->nparam), lp(nwg);
    TR_HIP(hipMemcpy(part.data(), t->gdb, part.size() * 4, hipMemcpyDeviceToHost));
    TR_HIP(hipMemcpy(lp.data(), t->gloss, lp.size() * 4, hipMemcpyDeviceToHost));
    if (grad_host)
        for (int p = 0; p < t->nparam; p++) {
            double s = 0.0;
            for (size_t w = 0; w < nwg; w++) s += part[w * t->nparam + p];
            grad_host[p] = (float)s;
        }
    if (loss_host) {
        double s = 0.0;
        for (size_t w = 0; w < nwg; w++) s += lp[w];
        *loss_host = s;
    }
    return WL_OK;
}

}  // extern "C"
