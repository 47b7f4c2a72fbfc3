/*
 * oracle.h -- CPU restatement of the reference's QC-LDPC codec.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (wimax_ldpc_lib_amd/)
 * links, loads or calls this code; only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg use it, to check results against.
 *
 * Every function restates a reference routine (paths relative to the
 * reference repository root):
 *   or_code_init      lib/ldpc.cc:11-63 (N, M), lib/ldpc.cc:69-1206 (table
 *                     choice) and the expansion python_ldpc/ldpc_utils.py:99-153
 *                     that generated the lib/matrices tables
 *   or_syndrome       lib/ldpc_decoder.cc:67-104
 *   or_decode_f32     lib/ldpc_decoder.cc:122-262 (NMS 0.815 as shipped, plus
 *                     the OMS / MS compile-time variants at :17-21,225-238,
 *                     a fixed-iteration flag, an iteration-count output and
 *                     the learned per-block offset variant)
 *   or_decode_i8      lib/ldpc_decoder.cc:108-118
 *   or_encode         lib/ldpc_encoder.cc:29-105
 *   or_decode_batch   lib/decode_example.cc:75-77,118-160 (OpenMP frame loop)
 *
 * Pinning (what checks each function independently of this file) is listed
 * in tests/test_oracle_pins.py; parity status per function:
 *   or_code_init   pinned: all 114 reference tables (oracle/_ref) + hashes
 *   or_encode      pinned: reference encoder build (oracle/_ref) + H*c = 0
 *   or_syndrome    pinned: brute-force XOR over dense H
 *   or_decode_f32  pinned: brute-force max-log check-node rule, exact-rational
 *                  rounding of the double step, an independent numpy layered
 *                  decoder, special cases (clean input, max_iter = 0, 10000
 *                  initialiser) and the iteration prefix property
 *   learned fwd/bwd pinned: finite differences + OMS special case
 *   or_code_init_qc pinned: equals or_code_init's tables for the base matrices
 *                  of all 114 802.16e codes; the decoder itself is table-driven
 */
#ifndef WL_ORACLE_H
#define WL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int rate;          /* coderate enum, include/enums.h:4-12 */
    int z;             /* lifting factor 24..96 step 4 */
    int n;             /* codeword length 24 z */
    int m;             /* parity checks */
    int k;             /* info bits n - m */
    int c;             /* max row degree (m_col_size) */
    int mb;            /* base rows (layers) */
    int d;             /* encoder constant D at this z */
    int x;             /* encoder constant X */
    int16_t* cn;       /* [m][c] variable indices, -1 padded (reference lib/matrices layout) */
    int* row_len;      /* [m] (lib/ldpc_decoder.cc:40-58) */
    int* blk;          /* [mb][c] base column of slot j of layer l, -1 padded */
    int* shift;        /* [mb][c] lifted circulant shift of slot j of layer l */
    int* deg;          /* [mb] layer degree */
} or_code;

enum { OR_NMS = 0, OR_OMS = 1, OR_MS = 2, OR_LEARNED = 3 };

typedef struct {
    int max_iter;
    int early_stop;          /* 1: as shipped (syndrome checks); 0: exactly max_iter iterations */
    int variant;             /* OR_NMS (as shipped), OR_OMS, OR_MS, OR_LEARNED */
    const float* offsets;    /* OR_LEARNED: [T][mb][c] offsets, iteration t uses min(t, T-1) */
    int n_offset_iters;      /* T */
} or_dec_cfg;

int  or_code_init(or_code* code, int rate, int z);
/* quasi-cyclic code from a base matrix [mb][nb] of shifts at lifting z */
int  or_code_init_qc(or_code* code, int nb, int mb, int z, const int16_t* base);
void or_code_free(or_code* code);

unsigned or_syndrome(const or_code* code, const int8_t* hd, int early_exit);

/* One check-row update on its own: L_in_post_out[0..len) holds the row's
 * variable values rx[v_n] and receives the posteriors, msg[0..len) the
 * previous message (in) and the new one (out).  beta: OR_LEARNED offsets. */
void or_row_update(int variant, int len, float* L_in_post_out, float* msg, const float* beta);

/* One frame.  rx (reference convention: >= 0 slices to bit 1) is replaced by
 * the posterior, hd receives n bytes of 0/1.  lmn is caller scratch of
 * m*c floats.  Returns what ldpc_decoder::decode returns; *iters_out gets the
 * number of iterations run (0 when the input already satisfies H). */
unsigned or_decode_f32(const or_code* code, const or_dec_cfg* cfg, float* rx,
                       int8_t* hd, float* lmn, int* iters_out);
unsigned or_decode_i8(const or_code* code, const or_dec_cfg* cfg, const int8_t* bits,
                      int8_t* hd, float* lmn, int* iters_out);

void or_encode(const or_code* code, const uint8_t* info, uint8_t* cw);
/* or_encode over nframes frames, OpenMP over frames; returns seconds */
double or_encode_batch(const or_code* code, const uint8_t* info, uint8_t* cw, size_t nframes, int nthreads);

/* Frame-parallel batch over nframes frames (OpenMP, one scratch per thread,
 * schedule(static)).  llr is [nframes][n] and is replaced by posteriors.
 * hd [nframes][n], ret/iters/synd_w [nframes] (nullable).  Returns seconds of
 * the timed decode region (steady clock), as lib/decode_example.cc:126-137. */
double or_decode_batch(const or_code* code, const or_dec_cfg* cfg, float* llr, size_t nframes,
                       int8_t* hd, uint8_t* ret, uint8_t* iters, uint16_t* synd_w, int nthreads);

/* Learned offset min-sum training reference: forward over T unrolled
 * iterations (no early stop) saving the per-row selection state, then the
 * exact backward of the mean binary cross-entropy over the n coded bits
 * (bit 1 <=> posterior > 0), or with per_iteration the mean over the T
 * iterations of the cross-entropy of each iteration's posteriors.
 * grad_off [T][mb][c] is accumulated (+=).  Returns the loss summed over
 * frames. */
double or_learned_fwd_bwd(const or_code* code, const float* offsets, int T,
                          const float* llr, const uint8_t* bits, size_t nframes,
                          float* post_out, float* grad_off, int per_iteration);

int or_omp_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
