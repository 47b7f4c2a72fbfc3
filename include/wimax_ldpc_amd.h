/*
 * wimax_ldpc_amd.h -- C ABI of the MI355X (gfx950) WiMAX QC-LDPC codec.
 *
 * The reference library (myersw12/wimax_ldpc_lib) has no C ABI: its public
 * interface is the C++ classes in include/ldpc.h, include/ldpc_decoder.h,
 * include/ldpc_encoder.h and include/BER_runner.h.  This header is the flat
 * boundary those classes (re-declared with the same signatures in this
 * directory) are implemented on; it takes plain pointers and sizes only.
 *
 * Conventions
 *   - Soft values use the reference convention: a value >= 0 (including -0.0)
 *     slices to bit 1 (lib/ldpc_decoder.cc:130,242; bit b is sent as 2b-1 at
 *     lib/BER_runner.cc:183).
 *   - Packed hard decisions: bit (v % 8) of byte (v / 8) is variable v.
 *   - Functions named *_dev take device pointers and enqueue on `stream`
 *     (a hipStream_t, NULL = the null stream); they do not synchronize.
 *     Functions named *_host take host pointers and return when done.
 *   - Every function returns 0 or a negative WL_E* code; wl_last_error()
 *     describes the last failure of the calling thread.  No C++ exception
 *     crosses this boundary.
 *   - A handle may be used by one host thread at a time (as the reference's
 *     objects: one decoder per thread, lib/decode_example.cc:75-77).
 */
#ifndef WIMAX_LDPC_AMD_H
#define WIMAX_LDPC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WL_OK        0
#define WL_EINVAL  (-22)   /* bad rate, z, size or pointer */
#define WL_ENOMEM  (-12)   /* device or host allocation failed */
#define WL_EDEVICE  (-5)   /* HIP runtime error (no device, launch failure, ...) */
#define WL_ENOTSUP (-95)   /* valid request the library does not support */

/* coderate values, identical to the reference's enum coderate (include/enums.h:4-12) */
#define WL_RATE_1_2   0
#define WL_RATE_2_3A  1
#define WL_RATE_2_3B  2
#define WL_RATE_3_4A  3
#define WL_RATE_3_4B  4
#define WL_RATE_5_6   5

/* decoder flags */
#define WL_FLAG_EARLY_STOP  0x1u  /* syndrome check before and after every iteration (as shipped) */
#define WL_FLAG_MIXED_UNORDERED 0x2u  /* wl_mixed: no reliability order inside code buckets */
/* message rule of wl_decoder (one at most; default: normalised min-sum, the
 * reference as shipped, lib/ldpc_decoder.cc:17-18,217-224) */
#define WL_FLAG_OMS 0x4u  /* offset min-sum s * max(min - 0.10, 0) in double: the
                             reference built with OMS defined (:20,225-231) */
#define WL_FLAG_MS  0x8u  /* plain min-sum s * min in float: the reference built
                             with neither NMS nor OMS (:234-237) */

typedef struct wl_decoder wl_decoder;
typedef struct wl_encoder wl_encoder;

typedef struct {
    int rate;
    int z;
    int n;     /* codeword bits, m_N (lib/ldpc.cc:16) */
    int m;     /* parity checks, m_M (lib/ldpc.cc:18-45) = m_row_size */
    int k;     /* info bits n - m */
    int c;     /* max row degree, m_col_size */
    int mb;    /* base rows = decoding layers */
    int d;     /* encoder constant m_D at this z */
    int x;     /* encoder constant m_X */
} wl_code_info;

/* Last error text of the calling thread ("" if none). */
const char* wl_last_error(void);

/* Library version string. */
const char* wl_version(void);

/* Number of visible gfx950 devices (0 when none; never fails). */
int wl_device_count(void);

/* Code parameters for (rate, z).  Replaces ldpc::ldpc (include/ldpc.h:55,
 * lib/ldpc.cc:11-63): WL_EINVAL where the reference throws. */
int wl_code_params(int rate, int z, wl_code_info* out);

/* Expanded parity-check table [m][c] with -1 padding, identical to the
 * reference's m_checknode_array (lib/ldpc.cc:69-1206, lib/matrices/). */
int wl_code_table(int rate, int z, int16_t* out, size_t out_len);

/* ---------------------------------------------------------------- decoder */

/* Replaces ldpc_decoder::ldpc_decoder(rate, z, max_iter, num_threads)
 * (include/ldpc_decoder.h:25-28, lib/ldpc_decoder.cc:27-59).  flags:
 * WL_FLAG_EARLY_STOP, and at most one of WL_FLAG_OMS / WL_FLAG_MS (the
 * offset and plain rules use code shape 2 at two frames per workgroup where
 * z % 32 == 0 and the register-lean shape otherwise, at 4 waves per SIMD,
 * whatever wl_decoder_set_variant chose; learned offsets, when set, replace
 * the rule, and the learned fallback re-decodes with normalised min-sum;
 * run-time QC codes: WL_ENOTSUP). */
int wl_decoder_create(int rate, int z, int max_iter, unsigned flags, wl_decoder** out);
int wl_decoder_destroy(wl_decoder* dec);

/* Launch shape of a decoder (no counterpart in the reference, whose decoder
 * is one frame per CPU thread).  frames_per_wg: frames decoded together by
 * one workgroup (0 = the autotuned launch shape, code shape included);
 * variant: code shape of the
 * normalised min-sum kernel (0: slot addresses held in registers, 1: slot
 * addresses recomputed per layer, fewer registers, 2: as 1 with each wave
 * holding 32 rows of two frames -- only for z % 32 == 0 and
 * frames_per_wg 2, 3: 16 rows of four frames -- only for z = 32, 48, 64 and
 * frames_per_wg 4; set frames_per_wg first, and another frames_per_wg then
 * switches 2 and 3 to 1; 4: as 0 compiled for four waves per SIMD, for
 * contiguous batches).  Results are identical for every shape; create()
 * starts from the autotuned table. */
int wl_decoder_set_frames_per_wg(wl_decoder* dec, int frames_per_wg);
int wl_decoder_get_frames_per_wg(const wl_decoder* dec, int* out);
int wl_decoder_set_variant(wl_decoder* dec, int variant);
int wl_decoder_get_variant(const wl_decoder* dec, int* out);

/* Small-batch shape (no counterpart in the reference): contiguous
 * normalised min-sum batches of at most max_frames frames are decoded one
 * frame per workgroup with two lanes per parity row, which halves the serial
 * work of a layer (batches that leave most of the GPU idle are bound by that
 * chain) and, with early stop, retires each frame's workgroup as soon as the
 * frame converges.  Results are identical to the other shapes.  0 = never;
 * WL_SMALL_BATCH_AUTO = the default for the current device: with early stop
 * up to 32 waves per SIMD (32768 frames at z <= 32 on 256 compute units,
 * 10922 at z = 96), without it up to 3 (3072 and 1024).  Learned, offset
 * and plain min-sum decoders and run-time codes do not use it (get reports
 * 0 for them). */
#define WL_SMALL_BATCH_AUTO ((size_t)-1)
int wl_decoder_set_small_batch(wl_decoder* dec, size_t max_frames);
int wl_decoder_get_small_batch(const wl_decoder* dec, size_t* max_frames);

/* Autotuned launch shape compiled into the library for (rate, z): returns 0,
 * or 1 when the table has no entry and the heuristic default is reported. */
int wl_tuned_launch(int rate, int z, int* frames_per_wg, int* variant);

/* Batch decode of device-resident frames.  Replaces the OpenMP frame loop of
 * lib/decode_example.cc:128-135 over ldpc_decoder::decode(float* rx,
 * int8_t* decoded) (include/ldpc_decoder.h:33, lib/ldpc_decoder.cc:122-262).
 *   llr      [nframes][n]  input, reference convention (not modified)
 *   post     [nframes][n]  posterior LLRs (what decode() leaves in rx), nullable
 *   hard     [nframes][n/8] packed hard decisions, nullable
 *   ret      [nframes]     decode()'s return value: 0 syndrome clean, 1 failed, nullable
 *   iters    [nframes]     iterations run (0 if the input already satisfies H), nullable
 *   synd     [nframes]     failing parity checks of the output hard decisions
 *                          (compute_syndrome(hd, false), lib/ldpc_decoder.cc:67-104), nullable
 * NaN LLRs are outside the bit-exact contract: the reference is compiled with
 * -ffast-math (lib/CMakeLists.txt:6), which leaves its comparisons with NaN
 * undefined.  A frame holding a NaN decodes to some result (its NaN
 * posteriors stay NaN); every other frame of the batch is unaffected. */
int wl_decode_f32_dev(wl_decoder* dec, const float* llr, size_t nframes, float* post,
                      uint8_t* hard, uint8_t* ret, uint8_t* iters, uint16_t* synd, void* stream);

/* Host-pointer batch with decode()'s exact semantics: rx [nframes][n] is
 * replaced by the posterior, decoded [nframes][n] receives 0/1 bytes, ret
 * [nframes] decode()'s return value, iters (nullable) the iteration count. */
int wl_decode_f32_host(wl_decoder* dec, float* rx, size_t nframes, int8_t* decoded,
                       uint8_t* ret, uint8_t* iters);

/* Hard-input adapter: ldpc_decoder::decode(int8_t* rx, int8_t* decoded)
 * (lib/ldpc_decoder.cc:108-118), soft = bit*2.0 - 1.0; rx is not modified. */
int wl_decode_i8_host(wl_decoder* dec, const int8_t* rx, size_t nframes, int8_t* decoded,
                      uint8_t* ret, uint8_t* iters);

/* ldpc_decoder::compute_syndrome(int8_t*, bool) (lib/ldpc_decoder.cc:67-104)
 * over one frame of n 0/1 bytes (host). */
int wl_compute_syndrome_host(const wl_decoder* dec, const int8_t* hd, int exit_early,
                             unsigned* out);

/* Learned offset min-sum mode: offsets [T][mb][c] (slot j of layer l is the
 * j-th nonzero block of base row l), T unrolled iterations; iteration t uses
 * offsets[min(t, T-1)].  msg = sign * max(mag - offset, 0) in FP32.  Passing
 * NULL returns the decoder to the reference's normalised min-sum. */
int wl_decoder_set_offsets(wl_decoder* dec, const float* offsets_host, int T);

/* Learned decode with fallback: frames whose hard decisions fail the
 * syndrome after the learned iterations are decoded again from their input
 * LLRs with normalised min-sum, fallback_iters iterations, early stop.
 * fallback (nullable) [nframes] gets 1 for re-decoded frames, 0 for the
 * others.  post must not overlap llr (WL_EINVAL): the failed frames are
 * decoded again from their input LLRs.  For z = 24 and
 * z = 96 this is one persistent kernel that re-queues failed frames in the
 * workgroup and decodes them itself; other z use a learned pass, a device-side
 * compaction of the failures and a normalised min-sum pass over them. */
int wl_decode_learned_fallback_dev(wl_decoder* dec, const float* llr, size_t nframes,
                                   int fallback_iters, float* post, uint8_t* hard,
                                   uint8_t* ret, uint8_t* iters, uint16_t* synd,
                                   uint8_t* fallback, void* stream);

/* Frames per call of wl_decode_learned_fallback_noalloc_dev. */
#define WL_NOALLOC_MAX_FRAMES (1 << 20)

/* wl_decode_learned_fallback_dev without any allocation (no counterpart in
 * the reference).  It is for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, and harnesses that
 * forbid hipMalloc.  Same arguments, kernels and results.  Its scratch is a
 * static device array (4 MiB + 16 B on every GPU the library runs on, held
 * until the process ends): the tile counter, and the list of failed frames
 * where z is not 24 or 96.  At most WL_NOALLOC_MAX_FRAMES frames per call
 * (WL_EINVAL beyond).  Calls on one GPU share that array, so calls that
 * overlap on one GPU (two streams, or two host threads) race; calls
 * serialised on one stream do not.  The decoder handle (offsets included)
 * is set up beforehand as for the other entry points. */
int wl_decode_learned_fallback_noalloc_dev(wl_decoder* dec, const float* llr, size_t nframes,
                                           int fallback_iters, float* post, uint8_t* hard,
                                           uint8_t* ret, uint8_t* iters, uint16_t* synd,
                                           uint8_t* fallback, void* stream);

/* ---------------------------------------------------------------- mixed batches */

/* Frames of different (rate, z) decoded together (no counterpart in the
 * reference, where one decoder object serves one code).  A frame's code id is
 * rate * 19 + (z - 24) / 4 (0..113).  Frame g holds n_g = 24 z_g LLRs starting
 * at element frame_off[g] of a ragged buffer; its posteriors go to the same
 * offsets of post, its packed hard decisions to byte frame_off[g] / 8 of hard,
 * and ret / iters / synd to index g.  A decode is one hipGraph launch: frames
 * are bucketed by code on the device and the buckets of all rates at one z
 * are decoded by one persistent kernel (each rate with its own code shape);
 * results equal a per-code wl_decode_f32_dev bit for bit.  With early stop (create flags),
 * frames inside a bucket are arranged by a reliability estimate (mean |LLR|
 * of the first 256 LLRs) so that the frames one workgroup decodes together
 * tend to stop together; WL_FLAG_MIXED_UNORDERED turns that off (same
 * results).  The graph is captured at the first call with a given set of
 * buffers and replayed afterwards.  Frames whose code id is invalid (or not
 * enabled, see wl_mixed_set_codes) are skipped. */
typedef struct wl_mixed wl_mixed;
/* flags: WL_FLAG_EARLY_STOP, WL_FLAG_MIXED_UNORDERED (WL_FLAG_OMS / MS:
 * WL_ENOTSUP, mixed batches decode with normalised min-sum) */
int wl_mixed_create(int max_iter, unsigned flags, wl_mixed** out);
int wl_mixed_destroy(wl_mixed* m);
/* Restrict the graph to the listed code ids (ncodes = 0: all 114). */
int wl_mixed_set_codes(wl_mixed* m, const int* codes, int ncodes);
/* Launch shape the graph uses for code (rate, z): frames per workgroup, code
 * shape variant and persistent grid of the kernel that decodes all rates of
 * that z. */
int wl_mixed_get_launch(const wl_mixed* m, int rate, int z, int* frames_per_wg, int* variant, int* grid);
/* Frames per workgroup of the kernel that decodes the codes of lifting size z
 * (0: the value the autotuner measured on mixed batches, compiled in).
 * Results do not depend on it. */
int wl_mixed_set_group_frames_per_wg(wl_mixed* m, int z, int frames_per_wg);
/* frame_off [nframes] = exclusive prefix sum of n over cfg [nframes] (device). */
int wl_mixed_offsets_dev(wl_mixed* m, const uint8_t* cfg, size_t nframes, uint64_t* frame_off, void* stream);
int wl_mixed_decode_dev(wl_mixed* m, const uint8_t* cfg, const uint64_t* frame_off, const float* llr,
                        size_t nframes, float* post, uint8_t* hard, uint8_t* ret, uint8_t* iters,
                        uint16_t* synd, void* stream);
/* Learned offsets [T][mb][c] for every code of `rate` (NULL clears). */
int wl_mixed_set_offsets(wl_mixed* m, int rate, const float* offsets_host, int T);
/* Learned decoding with re-queue: codes whose rate has offsets (and whose z is
 * 24 or 96) are decoded by one persistent kernel per code that runs the
 * learned decoder (max_iter iterations, the early-stop flag of create()) and
 * re-queues each frame whose syndrome still fails to normalised min-sum with
 * fallback_iters iterations and early stop, from the frame's input LLRs, in
 * the same kernel; fallback [nframes] (nullable) gets 1 for such frames
 * and 0 for every other frame.  post may not be llr (WL_EINVAL).
 * Other codes are decoded with normalised min-sum as wl_mixed_decode_dev. */
int wl_mixed_decode_learned_fallback_dev(wl_mixed* m, const uint8_t* cfg, const uint64_t* frame_off,
                                         const float* llr, size_t nframes, int fallback_iters, float* post,
                                         uint8_t* hard, uint8_t* ret, uint8_t* iters, uint16_t* synd,
                                         uint8_t* fallback, void* stream);
/* Frames skipped by the last decode for an invalid or disabled code id
 * (synchronizes the device). */
int wl_mixed_invalid_count(wl_mixed* m, uint64_t* out);
/* The code buckets the last decode of the handle used (no counterpart in the
 * reference; for checking the bucketing): buckets_host [2 * 114] gets
 * {first slot, frame count} per code id, and idx_dev (device, nullable)
 * [nframes] the frame ids in slot order, i.e. each code's frames from its
 * first slot on.  nframes must not exceed the last decode's batch.
 * Synchronizes the device. */
int wl_mixed_bucket_list(wl_mixed* m, uint32_t* buckets_host, uint32_t* idx_dev, size_t nframes);
/* Seeded synthetic mixed frames: frame g of code cfg[g] is exactly frame g of
 * wl_channel_dev on an encoder of that code (same seed). */
int wl_mixed_channel_dev(wl_mixed* m, uint64_t seed, const uint8_t* cfg, const uint64_t* frame_off,
                         size_t nframes, float ebn0_lo, float ebn0_hi, float* llr, uint8_t* bits, float* ebn0,
                         void* stream);

/* ---------------------------------------------------------------- trainer */

/* Learned offset min-sum trainer.  No counterpart in the reference, whose
 * normalisation / offset constants are fixed by hand (lib/ldpc_decoder.cc:17-21).
 * Model: T unrolled layered iterations with messages s * max(mag - beta, 0),
 * one offset per nonzero base-matrix block per iteration (offsets [T][mb][c],
 * the layout of wl_decoder_set_offsets); loss: mean binary cross-entropy of
 * the final posteriors over the n coded bits.  Each step draws `batch` fresh
 * frames on the GPU (wl_channel_dev's generator, seed `seed`, Eb/N0 uniform in
 * [ebn0_lo, ebn0_hi) per frame), runs the forward and backward kernels and an
 * Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8).  Kernels exist for z = 24
 * and z = 96 (WL_ENOTSUP otherwise) and for the five rates with a valid
 * encoder (3/4B: WL_ENOTSUP).  All offsets start at init_offset. */
typedef struct wl_trainer wl_trainer;
int wl_trainer_create(int rate, int z, int T, int batch, uint64_t seed, float lr, float init_offset,
                      wl_trainer** out);
int wl_trainer_destroy(wl_trainer* tr);
int wl_trainer_set_hparams(wl_trainer* tr, float lr, float beta1, float beta2, float eps);
/* per_iteration = 1: the loss is the mean over the T unrolled iterations of
 * the cross-entropy of each iteration's posteriors (each with weight 1/T);
 * 0 (default): the cross-entropy of the final posteriors.  Applies to
 * wl_trainer_step and wl_trainer_grad_dev. */
int wl_trainer_set_loss(wl_trainer* tr, int per_iteration);
/* nsteps training steps (one captured hipGraph launch each); last_loss
 * (nullable) receives the mean loss of the last step's batch. */
int wl_trainer_step(wl_trainer* tr, int nsteps, float ebn0_lo, float ebn0_hi, float* last_loss);
int wl_trainer_num_params(const wl_trainer* tr, int* T, int* mb, int* c);
int wl_trainer_get_offsets(wl_trainer* tr, float* offsets_host);
int wl_trainer_set_offsets(wl_trainer* tr, const float* offsets_host);
int wl_trainer_get_step(wl_trainer* tr, uint64_t* step);
/* Checkpoint / resume: offsets, Adam moments, step counter and seed. */
int wl_trainer_save(wl_trainer* tr, const char* path);
int wl_trainer_load(wl_trainer* tr, const char* path);
/* Loss (summed over frames and bits) and its gradient with respect to the
 * current offsets on caller frames: llr [nframes][n] and bits [nframes][n]
 * on the device; post (nullable, device) gets the final posteriors;
 * grad_host [T][mb][c] is the gradient of the MEAN loss. */
int wl_trainer_grad_dev(wl_trainer* tr, const float* llr, const uint8_t* bits, size_t nframes, float* post,
                        float* grad_host, double* loss_host);

/* ---------------------------------------------------------------- encoder */

/* Replaces ldpc_encoder::ldpc_encoder(rate, z, num_threads)
 * (include/ldpc_encoder.h:21, lib/ldpc_encoder.cc:9-17). */
int wl_encoder_create(int rate, int z, wl_encoder** out);
int wl_encoder_destroy(wl_encoder* enc);

/* An encoder that produces valid codewords for every rate, 3/4B included (no
 * counterpart in the reference, which cannot encode 3/4B: README.md:25, and
 * whose encoder gives codewords that fail H there).  It solves the parity
 * part by the general dual-diagonal back-substitution: the sum of all check
 * rows gives P^h_mid p0 = XOR of the lambda blocks, then p_1, p_2, ... from
 * the staircase.  Since the parity part is invertible, for the other five
 * rates its codewords equal wl_encoder_create's bit for bit.  All wl_encode_*
 * and wl_channel_dev calls accept it. */
int wl_encoder_create_standard(int rate, int z, wl_encoder** out);

/* ldpc_encoder::encode_data(uint8_t* info, uint8_t* codeword)
 * (include/ldpc_encoder.h:24, lib/ldpc_encoder.cc:50-105) over nframes
 * device-resident frames: info [nframes][k] bytes (nonzero = 1; copied as is
 * into the systematic part), codeword [nframes][n] bytes. */
int wl_encode_u8_dev(wl_encoder* enc, const uint8_t* info, size_t nframes, uint8_t* codeword,
                     void* stream);
int wl_encode_u8_host(wl_encoder* enc, const uint8_t* info, size_t nframes, uint8_t* codeword);

/* ---------------------------------------------------------------- channel */

/* Seeded synthetic frames on the device, the GPU counterpart of
 * BER_runner::run_iteration's source and channel (lib/BER_runner.cc:162-191):
 * uniform random info bits (Philox4x32-10 keyed by seed; counter = absolute
 * frame index and position), dual-diagonal encoding, BPSK bit 0 -> +1, AWGN
 * with sigma^2 = 1/(2 R 10^(EbN0/10)), R = k/n, LLR = 2y/sigma^2, written
 * negated, i.e. in the reference convention.  Frame f (absolute index
 * frame_offset + f) uses Eb/N0 = ebn0_lo when ebn0_hi <= ebn0_lo, else a
 * uniform draw in [ebn0_lo, ebn0_hi) from the same generator.
 *   llr  [nframes][n] floats (nullable)
 *   bits [nframes][n] codeword bytes 0/1 (nullable)
 *   ebn0 [nframes] the Eb/N0 used per frame (nullable) */
int wl_channel_dev(wl_encoder* enc, uint64_t seed, uint64_t frame_offset, size_t nframes,
                   float ebn0_lo, float ebn0_hi, float* llr, uint8_t* bits, float* ebn0,
                   void* stream);

/* ------------------------------------------------------- run-time QC codes */

/* Quasi-cyclic LDPC codes given at run time, beyond the 114 compiled 802.16e
 * tables: a base matrix of circulant shifts (802.11n-style tables, other
 * lifting sizes) or an alist file in the format the reference writes and
 * reads (python_ldpc/ldpc_utils.py:157-240, python_ldpc/ldpc_decoder.py:20-59).
 * Row r of base row l checks variable b z + (r + base[l][b]) mod z for every
 * block column b with base[l][b] >= 0.  Limits: mb <= 64 base rows, <= 32
 * blocks per row, z <= 512, the frame's posteriors and messages within 64 KiB. */
typedef struct wl_qc_code wl_qc_code;

/* base: [mb][nb] shifts at lifting z, -1 = zero block. */
int wl_qc_create(int nb, int mb, int z, const int16_t* base, wl_qc_code** out);

/* text[0..len): an alist file.  z > 0: that lifting size (the rows must form
 * z x z circulant blocks in layer-major order); z = 0: the largest one that
 * fits (z = 1 for a code with no such structure: every row its own layer). */
int wl_qc_from_alist(const char* text, size_t len, int z, wl_qc_code** out);
int wl_qc_destroy(wl_qc_code* q);

/* rate = -1; d and x are the dual-diagonal encoder constants, -1 when the
 * parity part does not have that structure. */
int wl_qc_params(const wl_qc_code* q, wl_code_info* out);
int wl_qc_base(const wl_qc_code* q, int* nb, int* mb, int16_t* out, size_t out_len);

/* expanded check-node table [m][c], ascending, -1 padded (the layout of
 * wl_code_table) */
int wl_qc_table(const wl_qc_code* q, int16_t* out, size_t out_len);

/* the code as alist text (reference writer's format); *len gets the length
 * without the terminating 0, out may be NULL to query it */
int wl_qc_alist(const wl_qc_code* q, char* out, size_t cap, size_t* len);

/* A decoder handle for the code: every wl_decode_* / wl_compute_syndrome_host
 * entry point then decodes it with the reference's normalised min-sum
 * (layered, one base row per layer), bit-exact as for the compiled codes.
 * Learned offsets are not available for run-time codes. */
int wl_decoder_create_qc(const wl_qc_code* q, int max_iter, unsigned flags, wl_decoder** out);

/* An encoder handle (wl_encode_u8_*, wl_channel_dev) for a code with the
 * dual-diagonal parity part (wl_qc_params: x >= 0); WL_ENOTSUP otherwise. */
int wl_encoder_create_qc(const wl_qc_code* q, wl_encoder** out);

#ifdef __cplusplus
}
#endif
#endif
