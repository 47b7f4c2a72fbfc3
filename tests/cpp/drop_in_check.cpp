// drop_in_check -- exercises the C++ drop-in classes (include/ldpc_decoder.h,
// ldpc_encoder.h, BER_runner.h) the way the reference's callers use them, and
// writes what they produce for tests/test_gpu_drop_in.py to check against
// the oracle.  Test infrastructure; needs a GPU.
//
//   drop_in_check frames <rate> <z> <max_iter> <threads> <llr.f32> <out_prefix>
//       one decoder object per OpenMP thread, decode(float*, int8_t*) per
//       frame (the pattern of lib/decode_example.cc:118-135), then the same
//       frames through decode_batch(); writes <out>.post/.dec/.ret and
//       <out>.bpost/.bdec/.bret
//   drop_in_check ber <rate> <z> <max_iter> <ebn0> <nframes> <out_prefix>
//       BER_runner::run_iteration nframes times, run_batch(nframes), and
//       run_batch in two parts; then restates the runner's frame source from
//       lib/BER_runner.cc:124-191 (LFSR data words, mt19937(0) +
//       normal_distribution<float> samples) and writes <out>.data (uint8
//       [F][k]) and <out>.noise (float [F][n]); prints one JSON line.
//   drop_in_check stream <rate> <z> <nframes> <out_prefix>
//       only the restated frame source: <out>.data and <out>.noise
//   drop_in_check until <rate> <z> <ebn0> <max_frames> <repeats>
//       the reference's "run_iteration until the first error" loop
//       (lib/Capacity_plot_example.cc:147-156) against
//       BER_runner::run_until_error, repeats times on one stream each;
//       prints one JSON line with both sequences
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

#include "BER_runner.h"
#include "ldpc_decoder.h"

using namespace wimax_ldpc_lib;

static std::vector<char> slurp(const char* path)
{
    std::vector<char> b;
    FILE* f = fopen(path, "rb");
    if (!f) return b;
    fseek(f, 0, SEEK_END);
    b.resize((size_t)ftell(f));
    fseek(f, 0, SEEK_SET);
    if (fread(b.data(), 1, b.size(), f) != b.size()) b.clear();
    fclose(f);
    return b;
}

static void spit(const std::string& path, const void* p, size_t bytes)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f || fwrite(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
}

static int frames_mode(char** a)
{
    const coderate rate = (coderate)atoi(a[0]);
    const unsigned z = atoi(a[1]), max_iter = atoi(a[2]), threads = atoi(a[3]);
    const size_t n = (size_t)(z / 96.0 * BASE_LDPC_BLOCK_LEN);
    std::vector<char> raw = slurp(a[4]);
    const size_t F = raw.size() / (n * sizeof(float));
    if (F == 0 || raw.size() != F * n * sizeof(float)) {
        fprintf(stderr, "bad llr file\n");
        return 2;
    }
    const std::string out = a[5];
    std::vector<float> rx(F * n), rx2(F * n);
    memcpy(rx.data(), raw.data(), raw.size());
    rx2 = rx;
    std::vector<ldpc_decoder*> decs(threads);
    for (unsigned t = 0; t < threads; t++) decs[t] = new ldpc_decoder(rate, z, max_iter, threads);
    std::vector<int8_t> dec(F * n), dec2(F * n);
    std::vector<uint8_t> ret(F), ret2(F);
    int bad = 0;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long f = 0; f < (long)F; f++) {
        try {
            ret[f] = (uint8_t)decs[omp_get_thread_num()]->decode(rx.data() + f * n, dec.data() + f * n);
        } catch (...) {
#pragma omp atomic write
            bad = 1;
        }
    }
    if (bad) {
        fprintf(stderr, "decode threw\n");
        return 3;
    }
    decs[0]->decode_batch(rx2.data(), F, dec2.data(), ret2.data(), nullptr);
    for (unsigned t = 0; t < threads; t++) delete decs[t];
    spit(out + ".post", rx.data(), rx.size() * 4);
    spit(out + ".dec", dec.data(), dec.size());
    spit(out + ".ret", ret.data(), ret.size());
    spit(out + ".bpost", rx2.data(), rx2.size() * 4);
    spit(out + ".bdec", dec2.data(), dec2.size());
    spit(out + ".bret", ret2.data(), ret2.size());
    printf("{\"frames\": %zu}\n", F);
    return 0;
}

// the frame source of lib/BER_runner.cc:124-191, restated: LFSR data words
// (x^31 + x^28 + 1, seed 0xdeadbeef, reset every 115200 bits) and the
// mt19937(0) / normal_distribution<float> samples of each frame
static void restated_source(unsigned n, unsigned k, unsigned F, std::vector<uint8_t>& data,
                            std::vector<float>& noise)
{
    uint32_t lfsr = 0xdeadbeef;
    unsigned lfsr_index = 0;
    std::mt19937 gen(0);
    std::normal_distribution<float> dist(0.0, 1.0);
    data.assign((size_t)F * k, 0);
    noise.assign((size_t)F * n, 0.0f);
    for (unsigned f = 0; f < F; f++) {
        for (unsigned i = 0; i < k; i++) {
            const uint32_t bit = ((lfsr >> 0) ^ (lfsr >> 3)) & 1;
            lfsr = (lfsr >> 1) | (bit << 30);
            if (++lfsr_index == 1152 * 100) {
                lfsr = 0xdeadbeef;
                lfsr_index = 0;
            }
            data[(size_t)f * k + i] = (uint8_t)bit;
        }
        for (unsigned j = 0; j < n; j++) noise[(size_t)f * n + j] = dist(gen);
    }
}

static int stream_mode(char** a)
{
    const coderate rate = (coderate)atoi(a[0]);
    const unsigned z = atoi(a[1]), F = atoi(a[2]);
    BER_runner r(rate, z, 1, 10);
    std::vector<uint8_t> data;
    std::vector<float> noise;
    restated_source(r.get_codeword_len(), r.get_dataword_len(), F, data, noise);
    const std::string out = a[3];
    spit(out + ".data", data.data(), data.size());
    spit(out + ".noise", noise.data(), noise.size() * 4);
    printf("{\"n\": %u, \"k\": %u, \"frames\": %u}\n", r.get_codeword_len(), r.get_dataword_len(), F);
    return 0;
}

static int until_mode(char** a)
{
    const coderate rate = (coderate)atoi(a[0]);
    const unsigned z = atoi(a[1]);
    const double ebno = atof(a[2]);
    const unsigned max_frames = atoi(a[3]), repeats = atoi(a[4]);
    BER_runner ra(rate, z, 1, 10), rb(rate, z, 1, 10);
    std::string ja = "[", jb = "[";
    for (unsigned r = 0; r < repeats; r++) {
        unsigned j = 0;
        double ber = 0.0;
        for (j = 0; j < max_frames; j++) {
            ber = ra.run_iteration(ebno);
            if (ber != 0.0) break;
        }
        const unsigned fa = ber != 0.0 ? j + 1 : max_frames;
        char buf[160];
        snprintf(buf, sizeof buf, "%s[%u, %d, %llu, %llu]", r ? ", " : "", fa, ber != 0.0 ? 1 : 0,
                 (unsigned long long)ra.get_total_num_bits(), (unsigned long long)ra.get_total_num_errors());
        ja += buf;
        ra.reset_test();
        size_t fb = 0;
        const bool err = rb.run_until_error(ebno, max_frames, &fb);
        snprintf(buf, sizeof buf, "%s[%zu, %d, %llu, %llu]", r ? ", " : "", fb, err ? 1 : 0,
                 (unsigned long long)rb.get_total_num_bits(), (unsigned long long)rb.get_total_num_errors());
        jb += buf;
        rb.reset_test();
    }
    // both streams continue from the same place: one more batch of frames
    const double tail_a = (ra.run_batch(ebno, 64), (double)ra.get_total_num_errors());
    const double tail_b = (rb.run_batch(ebno, 64), (double)rb.get_total_num_errors());
    printf("{\"loop\": %s], \"until\": %s], \"tail_loop\": %.0f, \"tail_until\": %.0f}\n", ja.c_str(), jb.c_str(),
           tail_a, tail_b);
    return 0;
}

static int ber_mode(char** a)
{
    const coderate rate = (coderate)atoi(a[0]);
    const unsigned z = atoi(a[1]), max_iter = atoi(a[2]);
    const double ebn0 = atof(a[3]);
    const unsigned F = atoi(a[4]);
    const std::string out = a[5];

    BER_runner ra(rate, z, 1, max_iter);
    for (unsigned f = 0; f < F; f++) ra.run_iteration(ebn0);
    BER_runner rb(rate, z, 1, max_iter);
    rb.run_batch(ebn0, F);
    BER_runner rc(rate, z, 1, max_iter);
    rc.run_batch(ebn0, F / 3);
    rc.run_batch(ebn0, F - F / 3);

    std::vector<uint8_t> data;
    std::vector<float> noise;
    restated_source(ra.get_codeword_len(), ra.get_dataword_len(), F, data, noise);
    const unsigned n = ra.get_codeword_len(), k = ra.get_dataword_len();
    spit(out + ".data", data.data(), data.size());
    spit(out + ".noise", noise.data(), noise.size() * 4);
    printf("{\"n\": %u, \"k\": %u, \"a_bits\": %llu, \"a_errors\": %llu, \"b_bits\": %llu, \"b_errors\": %llu, "
           "\"c_bits\": %llu, \"c_errors\": %llu}\n",
           n, k, (unsigned long long)ra.get_total_num_bits(), (unsigned long long)ra.get_total_num_errors(),
           (unsigned long long)rb.get_total_num_bits(), (unsigned long long)rb.get_total_num_errors(),
           (unsigned long long)rc.get_total_num_bits(), (unsigned long long)rc.get_total_num_errors());
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 8 && !strcmp(argv[1], "frames")) return frames_mode(argv + 2);
    if (argc == 8 && !strcmp(argv[1], "ber")) return ber_mode(argv + 2);
    if (argc == 6 && !strcmp(argv[1], "stream")) return stream_mode(argv + 2);
    if (argc == 7 && !strcmp(argv[1], "until")) return until_mode(argv + 2);
    fprintf(stderr, "usage: drop_in_check frames|ber ...\n");
    return 2;
}
