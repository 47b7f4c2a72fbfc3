/*
 * oracle.c -- CPU restatement of the reference codec (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
 * this file, to check results against; the product never calls it.  Build flags keep
 * IEEE semantics (no fast-math, no FMA contraction) so that the double-
 * precision message step rounds exactly as the reference's C++ expression.
 * Reference paths are relative to the reference repository root.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../wimax_ldpc_lib_amd/csrc/base_matrices.h"

#define OR_MAXDEG 32
#define OR_ENC_STACK 2048   /* parity checks of the largest 802.16e code: 1152 */

/* Lifting rule of python_ldpc/ldpc_utils.py:99-110: floor(p z / 96) for every
 * rate except 2/3A, which uses p mod z; p <= 0 is kept as is. */
static int lift(int rate, int p, int z)
{
    if (p <= 0) return p;
    if (rate == 1) return p % z;
    return (p * z) / WL_Z0;
}

int or_code_init(or_code* code, int rate, int z)
{
    memset(code, 0, sizeof(*code));
    if (rate < 0 || rate > 5) return -1;          /* lib/ldpc.cc:46-50 */
    if (z < 24 || z > 96 || (z % 4) != 0) return -2; /* lib/ldpc.cc:1198-1203 */
    const struct wl_base_info* bi = &wl_base_table[rate];
    code->rate = rate;
    code->z = z;
    code->n = 24 * z;                              /* (z/96.0)*2304, lib/ldpc.cc:16 */
    code->mb = bi->mb;
    code->m = bi->mb * z;                          /* lib/ldpc.cc:18-45 */
    code->k = code->n - code->m;
    code->x = bi->x;
    code->d = lift(rate, bi->d0, z);
    int c = 0;
    code->deg = (int*)calloc(bi->mb, sizeof(int));
    for (int l = 0; l < bi->mb; l++) {
        int dg = 0;
        for (int b = 0; b < WL_NB; b++) dg += bi->p[l][b] >= 0;
        code->deg[l] = dg;
        if (dg > c) c = dg;
    }
    code->c = c;
    code->blk = (int*)malloc(sizeof(int) * bi->mb * c);
    code->shift = (int*)malloc(sizeof(int) * bi->mb * c);
    code->cn = (int16_t*)malloc(sizeof(int16_t) * code->m * c);
    code->row_len = (int*)malloc(sizeof(int) * code->m);
    for (int l = 0; l < bi->mb; l++) {
        int j = 0;
        for (int b = 0; b < WL_NB; b++) {
            int p = bi->p[l][b];
            if (p < 0) continue;
            code->blk[l * c + j] = b;
            code->shift[l * c + j] = lift(rate, p, z);
            j++;
        }
        for (; j < c; j++) { code->blk[l * c + j] = -1; code->shift[l * c + j] = -1; }
        /* expanded rows: row r of layer l touches b z + (r + s) mod z, in
         * ascending column order (np.roll of the identity, ldpc_utils.py:127-136) */
        for (int r = 0; r < z; r++) {
            int m = l * z + r;
            for (j = 0; j < c; j++) {
                if (j < code->deg[l]) {
                    int b = code->blk[l * c + j], s = code->shift[l * c + j];
                    code->cn[m * c + j] = (int16_t)(b * z + (r + s) % z);
                } else {
                    code->cn[m * c + j] = -1;
                }
            }
            code->row_len[m] = code->deg[l];
        }
    }
    return 0;
}

/* A quasi-cyclic code given by its base matrix at lifting z (shift -1 = zero
 * block): the same expansion as or_code_init, rows layer-major, each row's
 * variables ascending -- the table the reference's table-driven decoder walks
 * (lib/ldpc_decoder.cc:143-240 reads only m_checknode_array and the row
 * lengths).  rate = -1; d = x = -1 (no encoder constants). */
int or_code_init_qc(or_code* code, int nb, int mb, int z, const int16_t* base)
{
    memset(code, 0, sizeof(*code));
    if (nb < 2 || mb < 1 || mb >= nb || z < 1 || !base) return -1;
    code->rate = -1;
    code->z = z;
    code->n = nb * z;
    code->mb = mb;
    code->m = mb * z;
    code->k = code->n - code->m;
    code->d = -1;
    code->x = -1;
    int c = 0;
    code->deg = (int*)calloc(mb, sizeof(int));
    for (int l = 0; l < mb; l++) {
        int dg = 0;
        for (int b = 0; b < nb; b++) dg += base[l * nb + b] >= 0;
        code->deg[l] = dg;
        if (dg > c) c = dg;
    }
    code->c = c;
    code->blk = (int*)malloc(sizeof(int) * mb * c);
    code->shift = (int*)malloc(sizeof(int) * mb * c);
    code->cn = (int16_t*)malloc(sizeof(int16_t) * code->m * c);
    code->row_len = (int*)malloc(sizeof(int) * code->m);
    for (int l = 0; l < mb; l++) {
        int j = 0;
        for (int b = 0; b < nb; b++) {
            if (base[l * nb + b] < 0) continue;
            code->blk[l * c + j] = b;
            code->shift[l * c + j] = base[l * nb + b];
            j++;
        }
        for (; j < c; j++) { code->blk[l * c + j] = -1; code->shift[l * c + j] = -1; }
        for (int r = 0; r < z; r++) {
            int m = l * z + r;
            for (j = 0; j < c; j++)
                code->cn[m * c + j] = (j < code->deg[l])
                    ? (int16_t)(code->blk[l * c + j] * z + (r + code->shift[l * c + j]) % z) : (int16_t)-1;
            code->row_len[m] = code->deg[l];
        }
    }
    return 0;
}

void or_code_free(or_code* code)
{
    free(code->cn); free(code->row_len); free(code->blk); free(code->shift); free(code->deg);
    memset(code, 0, sizeof(*code));
}

/* lib/ldpc_decoder.cc:67-104 */
unsigned or_syndrome(const or_code* code, const int8_t* hd, int early_exit)
{
    unsigned num_errors = 0;
    for (int i = 0; i < code->m; i++) {
        int8_t xorsum = 0;
        for (int k = 0; k < code->c; k++) {
            int16_t idx = code->cn[i * code->c + k];
            if (idx != -1) xorsum ^= hd[idx];
        }
        if (xorsum != 0) {
            if (early_exit) return 1;
            num_errors += 1;
        }
    }
    return num_errors;
}

/* volk_32f_binary_slicer_8i contract: out = (in >= 0) ? 1 : 0 (-0.0 -> 1).
 * Called at lib/ldpc_decoder.cc:130,242. */
static void slice(int8_t* hd, const float* rx, int n)
{
    for (int i = 0; i < n; i++) hd[i] = (rx[i] >= 0.0f) ? 1 : 0;
}

/* One layered sweep over all rows: lib/ldpc_decoder.cc:143-240.  The
 * reference keeps (max_iter+1) message slots and iteration i reads slot i-1
 * and writes slot i; a row's slot i-1 is dead once read, so one slot per row
 * updated in place gives identical values.  it0 is the 0-based iteration
 * (learned offsets only). */
/* One row of the layered sweep: lib/ldpc_decoder.cc:146-239.  rx is read
 * and written through the row's variable list cn[0..len), msg is the row's
 * message slot, beta the row's learned offsets (OR_LEARNED only). */
static inline __attribute__((always_inline)) void row_update(int variant, const int16_t* cn, int len, float* rx,
                                                            float* msg, const float* beta)
{
    float LNM[OR_MAXDEG];
    int sign = 1;                       /* bool sign = true */
    float first_minimum = 10000.0f;
    unsigned first_min_index = 0;
    for (int n = 0; n < len; n++) {
        LNM[n] = rx[cn[n]] - msg[n];
        sign ^= !signbit(LNM[n]);
        float lnm_abs = fabsf(LNM[n]);
        if (lnm_abs < first_minimum) {
            first_minimum = lnm_abs;
            first_min_index = (unsigned)n;
        }
    }
    for (int n = 0; n < len; n++) {
        float iter_sign = (float)((sign ^ !signbit(LNM[n])) * -2 + 1);
        float minimum;
        if ((unsigned)n == first_min_index) {
            minimum = 10000.0f;
            for (int d = 0; d < len; d++) {
                if ((unsigned)d != first_min_index) {
                    float a = fabsf(LNM[d]);
                    if (a < minimum) minimum = a;
                }
            }
        } else {
            minimum = first_minimum;
        }
        switch (variant) {
        case OR_NMS: {   /* :218-223, float * double -> double, one rounding */
            double v = (double)iter_sign * ((double)minimum * 0.815);
            msg[n] = (float)v;
            rx[cn[n]] = (float)((double)LNM[n] + v);
            break;
        }
        case OR_OMS: {   /* :227-231 */
            double t = (double)minimum - 0.10;
            double v = (double)iter_sign * (t > 0.0 ? t : 0.0);
            msg[n] = (float)v;
            rx[cn[n]] = (float)((double)LNM[n] + v);
            break;
        }
        case OR_MS: {    /* :236-237 */
            float v = iter_sign * minimum;
            msg[n] = v;
            rx[cn[n]] = LNM[n] + v;
            break;
        }
        default: {       /* learned offset, FP32: s * max(mag - beta, 0) */
            float t = minimum - beta[n];
            float v = iter_sign * (t > 0.0f ? t : 0.0f);
            msg[n] = v;
            rx[cn[n]] = LNM[n] + v;
            break;
        }
        }
    }
}

void or_row_update(int variant, int len, float* L_in_post_out, float* msg, const float* beta)
{
    int16_t cn[OR_MAXDEG];
    for (int n = 0; n < len; n++) cn[n] = (int16_t)n;
    row_update(variant, cn, len, L_in_post_out, msg, beta);
}

/* One layered sweep over all rows: lib/ldpc_decoder.cc:143-240.  The
 * reference keeps (max_iter+1) message slots and iteration i reads slot i-1
 * and writes slot i; a row's slot i-1 is dead once read, so one slot per row
 * updated in place gives identical values.  it0 is the 0-based iteration
 * (learned offsets only). */
static void sweep(const or_code* code, const or_dec_cfg* cfg, float* rx, float* lmn, int it0)
{
    const int c = code->c;
    /* the variant is a constant in each loop, so the row update is compiled
     * once per variant, as the reference's #ifdef selects one at build time */
    switch (cfg->variant) {
    case OR_NMS:
        for (int m = 0; m < code->m; m++)
            row_update(OR_NMS, code->cn + (size_t)m * c, code->row_len[m], rx, lmn + (size_t)m * c, NULL);
        break;
    case OR_OMS:
        for (int m = 0; m < code->m; m++)
            row_update(OR_OMS, code->cn + (size_t)m * c, code->row_len[m], rx, lmn + (size_t)m * c, NULL);
        break;
    case OR_MS:
        for (int m = 0; m < code->m; m++)
            row_update(OR_MS, code->cn + (size_t)m * c, code->row_len[m], rx, lmn + (size_t)m * c, NULL);
        break;
    default: {
        const int t = it0 < cfg->n_offset_iters ? it0 : cfg->n_offset_iters - 1;
        for (int m = 0; m < code->m; m++) {
            const float* beta = cfg->offsets + ((size_t)t * code->mb + (size_t)(m / code->z)) * c;
            row_update(OR_LEARNED, code->cn + (size_t)m * c, code->row_len[m], rx, lmn + (size_t)m * c, beta);
        }
        break;
    }
    }
}

/* lib/ldpc_decoder.cc:122-262 with the fixed-iteration flag. */
unsigned or_decode_f32(const or_code* code, const or_dec_cfg* cfg, float* rx,
                       int8_t* hd, float* lmn, int* iters_out)
{
    unsigned num_errors = 0;
    memset(lmn, 0, sizeof(float) * (size_t)code->m * code->c);  /* slot 0 is zero, :38 */
    slice(hd, rx, code->n);
    if (cfg->early_stop) {
        if (or_syndrome(code, hd, 1) == 0) {
            if (iters_out) *iters_out = 0;
            return 0;
        }
    }
    for (int i = 1; i < cfg->max_iter + 1; i++) {
        sweep(code, cfg, rx, lmn, i - 1);
        slice(hd, rx, code->n);
        if (cfg->early_stop) {
            num_errors = or_syndrome(code, hd, 1);
            if (num_errors == 0) {
                if (iters_out) *iters_out = i;
                return 0;
            }
        }
    }
    if (iters_out) *iters_out = cfg->max_iter;
    if (!cfg->early_stop) num_errors = or_syndrome(code, hd, 1);
    return num_errors;
}

/* lib/ldpc_decoder.cc:108-118: soft = bit*2.0 - 1.0 */
unsigned or_decode_i8(const or_code* code, const or_dec_cfg* cfg, const int8_t* bits,
                      int8_t* hd, float* lmn, int* iters_out)
{
    float* soft = (float*)malloc(sizeof(float) * code->n);
    for (int i = 0; i < code->n; i++) soft[i] = (float)(bits[i] * 2.0 - 1.0);
    unsigned r = or_decode_f32(code, cfg, soft, hd, lmn, iters_out);
    free(soft);
    return r;
}

/* lib/ldpc_encoder.cc:29-105 (Lin-Wei-Ku parity prediction / correction). */
void or_encode(const or_code* code, const uint8_t* info, uint8_t* cw)
{
    const int M = code->m, z = code->z, X = code->x, D = code->d, C = code->c;
    const int K = code->n - code->m;
    /* the reference's member buffers m_P, m_V (include/ldpc_encoder.h:15-16),
     * cleared per frame (:55-56) */
    uint8_t Pb[OR_ENC_STACK], Vb[OR_ENC_STACK], Yb[OR_ENC_STACK];
    const int big = M > OR_ENC_STACK || z > OR_ENC_STACK;   /* run-time codes beyond 802.16e sizes */
    uint8_t* P = big ? (uint8_t*)malloc((size_t)M) : Pb;
    uint8_t* V = big ? (uint8_t*)malloc((size_t)M) : Vb;
    uint8_t* Y = big ? (uint8_t*)malloc((size_t)z) : Yb;
    memset(P, 0, (size_t)M);
    memset(V, 0, (size_t)M);
    memset(Y, 0, (size_t)z);
    for (int m = 0; m < M; m++) {                 /* compute_v, :35-47 */
        uint8_t value = 0;
        for (int n = 0; n < C; n++) {
            int16_t idx = code->cn[m * C + n];
            if (idx != -1 && (uint16_t)idx < K) value ^= (info[idx] != 0);
        }
        V[m] = value;
    }
    for (int i = M - 1; i > M - 1 - z; i--) P[i] = V[i] ^ P[M - 1 - i];      /* :61-64 */
    for (int i = z; i < (X + 1) * z; i++) P[i] = V[i - z] ^ P[i - z];       /* :67-70 */
    for (int i = M - 1 - z; i > (X + 1) * z - 1; i--) P[i] = V[i] ^ P[i + z]; /* :73-76 */
    for (int i = X * z; i < X * z + z; i++) Y[i - X * z] = V[i] ^ P[z + i];  /* :79-82 */
    for (int i = 0; i < z; i++) P[i] = Y[i] ^ P[z * X + i];                 /* :85-88 */
    memcpy(cw, info, (size_t)K);                                            /* :90 */
    for (int i = 0; i < M; i++)                                             /* :93-103 */
        cw[K + i] = (i < z) ? P[i] : (uint8_t)(P[i] ^ P[(i + D) % z]);
    if (big) {
        free(P);
        free(V);
        free(Y);
    }
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int or_omp_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* lib/decode_example.cc:75-77,118-160: one decoder state per thread and an
 * OpenMP frame loop; only the parallel region is timed. */
double or_decode_batch(const or_code* code, const or_dec_cfg* cfg, float* llr, size_t nframes,
                       int8_t* hd, uint8_t* ret, uint8_t* iters, uint16_t* synd_w, int nthreads)
{
    if (nthreads <= 0) nthreads = or_omp_max_threads();
    const size_t n = (size_t)code->n;
    double t0 = now_s();
#pragma omp parallel num_threads(nthreads)
    {
        float* lmn = (float*)malloc(sizeof(float) * (size_t)code->m * code->c);
        int8_t* hloc = (int8_t*)malloc(n);
#pragma omp for schedule(static)
        for (long f = 0; f < (long)nframes; f++) {
            int it = 0;
            int8_t* h = hd ? hd + (size_t)f * n : hloc;
            unsigned r = or_decode_f32(code, cfg, llr + (size_t)f * n, h, lmn, &it);
            if (ret) ret[f] = (uint8_t)r;
            if (iters) iters[f] = (uint8_t)it;
            if (synd_w) synd_w[f] = (uint16_t)or_syndrome(code, h, 0);
        }
        free(lmn);
        free(hloc);
    }
    return now_s() - t0;
}

/* A frame-parallel batch of or_encode (OpenMP over frames; the reference
 * encodes one frame per call, lib/encode_example.cc:131-158, with an OpenMP
 * region inside compute_v, lib/ldpc_encoder.cc:34).  Returns the seconds of
 * the timed region. */
double or_encode_batch(const or_code* code, const uint8_t* info, uint8_t* cw, size_t nframes, int nthreads)
{
    if (nthreads <= 0) nthreads = or_omp_max_threads();
    const size_t n = (size_t)code->n, k = (size_t)(code->n - code->m);
    double t0 = now_s();
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long f = 0; f < (long)nframes; f++) or_encode(code, info + (size_t)f * k, cw + (size_t)f * n);
    return now_s() - t0;
}

/* ---------------- learned offset min-sum: forward + exact backward -------- */

typedef struct {
    float min1, min2;
    int idx1, idx2;      /* -1 when the minimum is the 10000 initialiser */
    uint32_t sb;         /* bit j = signbit(L_j) */
    int parity;          /* 'sign' after the first pass */
} or_rowstate;

/* the cross-entropy term of posteriors P (bit 1 <=> P > 0) with weight w:
 * adds w * (softplus(P) - b P) to *loss, and w * (sigmoid(P) - b) * scale to
 * gP (or stores it, set = 1) */
static void bce_term(const float* P, const uint8_t* bits, int n, double w, double scale, double* loss, float* gP,
                     int set)
{
    for (int v = 0; v < n; v++) {
        double x = P[v], b = bits[v] ? 1.0 : 0.0;
        double sp = (x > 0 ? x : 0) + log1p(exp(-fabs(x)));
        *loss += w * (sp - b * x);
        double sg = 1.0 / (1.0 + exp(-x));
        float g = (float)((sg - b) * scale * w);
        gP[v] = set ? g : gP[v] + g;
    }
}

double or_learned_fwd_bwd(const or_code* code, const float* offsets, int T,
                          const float* llr, const uint8_t* bits, size_t nframes,
                          float* post_out, float* grad_off, int per_iteration)
{
    const int c = code->c, M = code->m, n = code->n, z = code->z;
    double loss = 0.0;
    /* loss weights: the final posteriors only, or 1/T for every iteration's */
    const int multi = per_iteration && T > 1;
    const double w_final = multi ? 1.0 / T : 1.0, w_it = multi ? 1.0 / T : 0.0;
    float* Pit = multi ? (float*)malloc(sizeof(float) * (size_t)(T - 1) * n) : NULL;
    float* P = (float*)malloc(sizeof(float) * n);
    float* Mmsg = (float*)malloc(sizeof(float) * (size_t)M * c);
    float* gP = (float*)malloc(sizeof(float) * n);
    float* gM = (float*)malloc(sizeof(float) * (size_t)M * c);
    or_rowstate* st = (or_rowstate*)malloc(sizeof(or_rowstate) * (size_t)T * M);
    double* gacc = (double*)calloc((size_t)T * code->mb * c, sizeof(double));
    const double scale = 1.0 / ((double)nframes * n);
    for (size_t f = 0; f < nframes; f++) {
        memcpy(P, llr + f * n, sizeof(float) * n);
        memset(Mmsg, 0, sizeof(float) * (size_t)M * c);
        for (int t = 0; t < T; t++) {
            for (int m = 0; m < M; m++) {
                const int len = code->row_len[m];
                const int16_t* cn = code->cn + (size_t)m * c;
                float* msg = Mmsg + (size_t)m * c;
                const float* beta = offsets + ((size_t)t * code->mb + m / z) * c;
                float L[OR_MAXDEG];
                int sign = 1, i1 = -1, i2 = -1;
                float mn1 = 10000.0f, mn2 = 10000.0f;
                uint32_t sb = 0;
                for (int j = 0; j < len; j++) {
                    L[j] = P[cn[j]] - msg[j];
                    sign ^= !signbit(L[j]);
                    sb |= (uint32_t)(signbit(L[j]) ? 1u : 0u) << j;
                    float a = fabsf(L[j]);
                    if (a < mn1) { mn1 = a; i1 = j; }
                }
                for (int j = 0; j < len; j++) {
                    if (j == i1 || (i1 < 0 && j == 0)) continue;
                    float a = fabsf(L[j]);
                    if (a < mn2) { mn2 = a; i2 = j; }
                }
                or_rowstate* s = &st[(size_t)t * M + m];
                s->min1 = mn1; s->min2 = mn2; s->idx1 = i1; s->idx2 = i2; s->sb = sb; s->parity = sign;
                int am = i1 < 0 ? 0 : i1;   /* the reference's first_min_index starts at 0 */
                for (int j = 0; j < len; j++) {
                    float isg = (float)((sign ^ !signbit(L[j])) * -2 + 1);
                    float mag = (j == am) ? mn2 : mn1;
                    float tt = mag - beta[j];
                    float v = isg * (tt > 0.0f ? tt : 0.0f);
                    msg[j] = v;
                    P[cn[j]] = L[j] + v;
                }
            }
            if (multi && t < T - 1) memcpy(Pit + (size_t)t * n, P, sizeof(float) * n);
        }
        if (post_out) memcpy(post_out + f * n, P, sizeof(float) * n);
        /* mean BCE over coded bits, bit 1 <=> P > 0: softplus(P) - b P */
        bce_term(P, bits + f * n, n, w_final, scale, &loss, gP, 1);
        memset(gM, 0, sizeof(float) * (size_t)M * c);
        for (int t = T - 1; t >= 0; t--) {
            /* the loss term of iteration t enters at its posteriors */
            if (multi && t < T - 1) bce_term(Pit + (size_t)t * n, bits + f * n, n, w_it, scale, &loss, gP, 0);
            for (int m = M - 1; m >= 0; m--) {
                const int len = code->row_len[m];
                const int16_t* cn = code->cn + (size_t)m * c;
                const int l = m / z;
                const float* beta = offsets + ((size_t)t * code->mb + l) * c;
                const or_rowstate* s = &st[(size_t)t * M + m];
                int am = s->idx1 < 0 ? 0 : s->idx1;
                float dL[OR_MAXDEG];
                float dmsg[OR_MAXDEG];
                for (int j = 0; j < len; j++) {
                    dmsg[j] = gP[cn[j]] + gM[(size_t)m * c + j];
                    dL[j] = gP[cn[j]];
                }
                for (int j = 0; j < len; j++) {
                    int sbj = (s->sb >> j) & 1;
                    int nonneg = !sbj;
                    float isg = (float)((s->parity ^ nonneg) * -2 + 1);
                    float mag = (j == am) ? s->min2 : s->min1;
                    if (!(mag - beta[j] > 0.0f)) continue;
                    float dmag = isg * dmsg[j];
                    gacc[((size_t)t * code->mb + l) * c + j] -= (double)dmag;
                    int k = (j == am) ? s->idx2 : s->idx1;
                    if (k < 0) continue;
                    float sk = ((s->sb >> k) & 1) ? -1.0f : 1.0f;
                    dL[k] += sk * dmag;
                }
                for (int j = 0; j < len; j++) {
                    gP[cn[j]] = dL[j];
                    gM[(size_t)m * c + j] = -dL[j];
                }
            }
        }
    }
    for (size_t i = 0; i < (size_t)T * code->mb * c; i++) grad_off[i] += (float)gacc[i];
    free(P); free(Mmsg); free(gP); free(gM); free(st); free(gacc); free(Pit);
    return loss;
}
