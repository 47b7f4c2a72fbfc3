// ldpc_classes.cpp -- the reference's C++ classes (include/ldpc.h,
// ldpc_decoder.h, ldpc_encoder.h, BER_runner.h) over the C ABI.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "BER_runner.h"
#include "ldpc.h"
#include "ldpc_decoder.h"
#include "ldpc_encoder.h"
#include "matrices.hpp"
#include "wimax_ldpc_amd.h"

namespace wimax_ldpc_lib {

namespace {
void check(int rc, const char* what)
{
    if (rc != WL_OK) throw std::runtime_error(std::string(what) + ": " + wl_last_error());
}
}  // namespace

// ----------------------------------------------------------------- ldpc

ldpc::ldpc()
    : m_checknode_array(nullptr), m_D(0), m_N(0), m_M(0), m_X(0), m_z(0), m_col_size(0), m_row_size(0),
      m_num_threads(0)
{
}

// lib/ldpc.cc:11-63
ldpc::ldpc(coderate rate, unsigned int z_factor, unsigned int num_threads) : ldpc()
{
    m_num_threads = num_threads;
    wl_code_info ci;
    if (wl_code_params((int)rate, (int)z_factor, &ci) != WL_OK) {
        printf("[!]ldpc - %s\n", wl_last_error());
        throw std::exception();
    }
    m_z = z_factor;
    m_N = ci.n;
    m_M = ci.m;
    m_D = ci.d;
    m_X = ci.x;
    m_row_size = ci.m;
    m_col_size = ci.c;
    // the code's table of matrices.h, as set_checknode_array (lib/ldpc.cc:69-1206)
    m_checknode_array = wl::matrix_table((int)rate, (int)z_factor);
    if (!m_checknode_array) throw std::runtime_error("no table for this code");
}

ldpc::~ldpc() {}

uint64_t ldpc::get_nanoseconds()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::high_resolution_clock::now().time_since_epoch())
        .count();
}

// ----------------------------------------------------------------- decoder

ldpc_decoder::ldpc_decoder() : m_max_iter(0), m_rate(HALFRATE), m_handle(nullptr), m_flags(WL_FLAG_EARLY_STOP) {}

ldpc_decoder::ldpc_decoder(coderate rate, unsigned int z_factor, unsigned int max_iter,
                           unsigned int num_threads)
    : ldpc(rate, z_factor, num_threads), m_max_iter(max_iter), m_rate(rate), m_handle(nullptr),
      m_flags(WL_FLAG_EARLY_STOP)
{
    check(wl_decoder_create((int)rate, (int)z_factor, (int)max_iter, m_flags, &m_handle), "wl_decoder_create");
}

ldpc_decoder::~ldpc_decoder()
{
    if (m_handle) wl_decoder_destroy(m_handle);
}

unsigned int ldpc_decoder::compute_syndrome(int8_t* rx_codeword, bool exit_early)
{
    unsigned out = 0;
    check(wl_compute_syndrome_host(m_handle, rx_codeword, exit_early ? 1 : 0, &out), "compute_syndrome");
    return out;
}

unsigned int ldpc_decoder::decode(int8_t* rx_codeword, int8_t* decoded)
{
    uint8_t ret = 0;
    check(wl_decode_i8_host(m_handle, rx_codeword, 1, decoded, &ret, nullptr), "decode");
    return ret;
}

unsigned int ldpc_decoder::decode(float* rx_codeword, int8_t* decoded)
{
    uint8_t ret = 0;
    check(wl_decode_f32_host(m_handle, rx_codeword, 1, decoded, &ret, nullptr), "decode");
    return ret;
}

void ldpc_decoder::decode_batch(float* rx, size_t nframes, int8_t* decoded, uint8_t* ret, uint8_t* iters)
{
    std::vector<uint8_t> tmp;
    if (!ret) {
        tmp.resize(nframes);
        ret = tmp.data();
    }
    check(wl_decode_f32_host(m_handle, rx, nframes, decoded, ret, iters), "decode_batch");
}

void ldpc_decoder::decode_batch(const int8_t* rx, size_t nframes, int8_t* decoded, uint8_t* ret,
                                uint8_t* iters)
{
    std::vector<uint8_t> tmp;
    if (!ret) {
        tmp.resize(nframes);
        ret = tmp.data();
    }
    check(wl_decode_i8_host(m_handle, rx, nframes, decoded, ret, iters), "decode_batch");
}

// a new handle with other flags; frames per workgroup carry over
void ldpc_decoder::recreate(unsigned int flags)
{
    if (!m_handle) throw std::runtime_error("decoder without a code");
    int f = 0;
    check(wl_decoder_get_frames_per_wg(m_handle, &f), "decoder flags");
    wl_decoder* h = nullptr;
    check(wl_decoder_create((int)m_rate, (int)m_z, (int)m_max_iter, flags, &h), "decoder flags");
    wl_decoder_set_frames_per_wg(h, f);
    wl_decoder_destroy(m_handle);
    m_handle = h;
    m_flags = flags;
}

void ldpc_decoder::set_early_stop(bool early_stop)
{
    recreate(early_stop ? (m_flags | WL_FLAG_EARLY_STOP) : (m_flags & ~WL_FLAG_EARLY_STOP));
}

void ldpc_decoder::set_message_rule(message_rule rule)
{
    const unsigned base = m_flags & ~(WL_FLAG_OMS | WL_FLAG_MS);
    recreate(rule == RULE_OMS ? (base | WL_FLAG_OMS) : (rule == RULE_MS ? (base | WL_FLAG_MS) : base));
}

// ----------------------------------------------------------------- encoder

ldpc_encoder::ldpc_encoder(coderate rate, unsigned int z_factor, unsigned int num_threads)
    : ldpc(rate, z_factor, num_threads), m_handle(nullptr)
{
    check(wl_encoder_create((int)rate, (int)z_factor, &m_handle), "wl_encoder_create");
}

ldpc_encoder::ldpc_encoder(coderate rate, unsigned int z_factor, unsigned int num_threads, bool standard)
    : ldpc(rate, z_factor, num_threads), m_handle(nullptr)
{
    if (standard)
        check(wl_encoder_create_standard((int)rate, (int)z_factor, &m_handle), "wl_encoder_create_standard");
    else
        check(wl_encoder_create((int)rate, (int)z_factor, &m_handle), "wl_encoder_create");
}

ldpc_encoder::~ldpc_encoder()
{
    if (m_handle) wl_encoder_destroy(m_handle);
}

void ldpc_encoder::encode_data(uint8_t* infoword, uint8_t* codeword)
{
    check(wl_encode_u8_host(m_handle, infoword, 1, codeword), "encode_data");
}

void ldpc_encoder::encode_batch(const uint8_t* infoword, size_t nframes, uint8_t* codeword)
{
    check(wl_encode_u8_host(m_handle, infoword, nframes, codeword), "encode_batch");
}

// ----------------------------------------------------------------- BER_runner

#define BITS_PER_SYM 1   // lib/BER_runner.cc:9

// lib/BER_runner.cc:13: a runner with no code.  The reference leaves every
// member unset; here the counters are zero and the frame methods throw.
BER_runner::BER_runner()
    : ldpc(), m_encoder(nullptr), m_decoder(nullptr), m_num_bits(0), m_num_errors(0), m_codeword_len(0),
      m_dataword_len(0), m_float_rate(0.0f), m_lfsr(0xdeadbeef), m_lfsr_start_state(0xdeadbeef), m_lfsr_index(0),
      m_lfsr_reset_len(1152 * 100)
{
}

void BER_runner::require_code() const
{
    if (!m_encoder || !m_decoder) {
        printf("[!]ber_runner - no code: constructed with BER_runner()\n");
        throw std::exception();
    }
}

// lib/BER_runner.cc:15-87
BER_runner::BER_runner(coderate rate, unsigned int z_factor, unsigned int num_threads, unsigned int max_iter)
    : ldpc(rate, z_factor, num_threads), m_encoder(nullptr), m_decoder(nullptr)
{
    m_num_bits = 0;
    m_num_errors = 0;
    m_codeword_len = (z_factor / 96.0) * BASE_LDPC_BLOCK_LEN;
    m_dist_awgn = std::normal_distribution<float>(0.0, 1.0);
    m_generator_awgn = std::mt19937(0);
    m_lfsr_start_state = 0xdeadbeef;
    m_lfsr = m_lfsr_start_state;
    m_lfsr_reset_len = 1152 * 100;
    m_lfsr_index = 0;
    switch (rate) {
    case HALFRATE:
        m_dataword_len = m_codeword_len / 2;
        m_float_rate = 0.5;
        break;
    case TWOTHIRDSA:
    case TWOTHIRDSB:
        m_dataword_len = (m_codeword_len * 2) / 3;
        m_float_rate = 2.0 / 3.0;
        break;
    case THREEQUARTERSA:
        m_dataword_len = (m_codeword_len * 3) / 4;
        m_float_rate = 3.0 / 4.0;
        break;
    case THREEQUARTERSB:
        printf("[!]ber_runner - 3/4B Not Implemented\n");
        throw std::exception();
    case FIVESIXTHS:
        m_dataword_len = (m_codeword_len * 5) / 6;
        m_float_rate = 5.0 / 6.0;
        break;
    default:
        printf("[!]test_encoder - Invalid Coderate: %d\n", rate);
        throw std::exception();
    }
    m_encoder = new ldpc_encoder(rate, z_factor, num_threads);
    m_decoder = new ldpc_decoder(rate, z_factor, max_iter, num_threads);
}

BER_runner::~BER_runner()
{
    delete m_encoder;
    delete m_decoder;
}

void BER_runner::reset_test()
{
    m_num_bits = 0;
    m_num_errors = 0;
}

uint64_t BER_runner::get_total_num_bits() { return m_num_bits; }
uint64_t BER_runner::get_total_num_errors() { return m_num_errors; }
unsigned int BER_runner::get_codeword_len() { return m_codeword_len; }
unsigned int BER_runner::get_dataword_len() { return m_dataword_len; }

// lib/BER_runner.cc:124-138: x^31 + x^28 + 1
uint32_t BER_runner::lfsr_get_num()
{
    uint32_t bit = ((m_lfsr >> 0) ^ (m_lfsr >> 3)) & 1;
    m_lfsr = (m_lfsr >> 1) | (bit << 30);
    m_lfsr_index++;
    if (m_lfsr_index == m_lfsr_reset_len) {
        m_lfsr = m_lfsr_start_state;
        m_lfsr_index = 0;
    }
    return bit;
}

void BER_runner::fill_with_random(uint8_t* buffer, unsigned int buf_len)
{
    for (unsigned int i = 0; i < buf_len; i++) buffer[i] = (uint8_t)lfsr_get_num();
}

// lib/BER_runner.cc:148-160
unsigned int BER_runner::compare_data(const uint8_t* original_data, const int8_t* decoded_data)
{
    unsigned int num_errors = 0;
    for (unsigned int j = 0; j < m_codeword_len; j++)
        if (original_data[j] != decoded_data[j]) num_errors++;
    return num_errors;
}

// BPSK and AWGN of lib/BER_runner.cc:172-191 for one encoded frame, with the
// same operand types (float sample + double sqrt(0.5) * float amplitude *
// float normal, rounded to float).
void BER_runner::make_frame(double EbNo_dB, uint8_t* dataword, uint8_t* codeword, float* channel)
{
    (void)dataword;
    double EbNo_linear = pow(10.0, (EbNo_dB / 10.0));
    float noise_amplitude = (1.0 / std::sqrt(EbNo_linear)) * std::sqrt(1.0 / (m_float_rate * BITS_PER_SYM));
    for (unsigned int k = 0; k < m_codeword_len; k++) channel[k] = codeword[k] * 2 - 1;
    for (unsigned int j = 0; j < m_codeword_len; j++)
        channel[j] = channel[j] + std::sqrt(0.5) * noise_amplitude * m_dist_awgn(m_generator_awgn);
}

// lib/BER_runner.cc:162-204
double BER_runner::run_iteration(double EbNo_dB)
{
    require_code();
    std::vector<uint8_t> data(m_dataword_len), cw(m_codeword_len);
    std::vector<float> ch(m_codeword_len);
    std::vector<int8_t> dec(m_codeword_len);
    fill_with_random(data.data(), m_dataword_len);
    m_encoder->encode_data(data.data(), cw.data());
    make_frame(EbNo_dB, data.data(), cw.data(), ch.data());
    m_decoder->decode(ch.data(), dec.data());
    m_num_bits += m_dataword_len;
    m_num_errors += compare_data(cw.data(), dec.data());
    return (double)m_num_errors / (double)m_num_bits;
}

double BER_runner::run_batch(double EbNo_dB, size_t nframes)
{
    require_code();
    if (nframes == 0) return m_num_bits ? (double)m_num_errors / (double)m_num_bits : 0.0;
    const size_t k = m_dataword_len, n = m_codeword_len;
    std::vector<uint8_t> data(nframes * k), cw(nframes * n);
    std::vector<float> ch(nframes * n);
    std::vector<int8_t> dec(nframes * n);
    for (size_t f = 0; f < nframes; f++) fill_with_random(data.data() + f * k, (unsigned)k);
    m_encoder->encode_batch(data.data(), nframes, cw.data());
    for (size_t f = 0; f < nframes; f++) make_frame(EbNo_dB, data.data() + f * k, cw.data() + f * n, ch.data() + f * n);
    m_decoder->decode_batch(ch.data(), nframes, dec.data(), nullptr, nullptr);
    for (size_t f = 0; f < nframes; f++) {
        m_num_bits += m_dataword_len;
        m_num_errors += compare_data(cw.data() + f * n, dec.data() + f * n);
    }
    return (double)m_num_errors / (double)m_num_bits;
}

bool BER_runner::run_until_error(double EbNo_dB, size_t max_frames, size_t* frames_run)
{
    require_code();
    const size_t k = m_dataword_len, n = m_codeword_len;
    size_t done = 0, chunk = 32;
    while (done < max_frames) {
        const size_t c = std::min(chunk, max_frames - done);
        // frame source state before the chunk, to rewind to an error frame
        const uint32_t lfsr = m_lfsr;
        const unsigned lfsr_index = m_lfsr_index;
        const std::mt19937 gen = m_generator_awgn;
        const std::normal_distribution<float> dist = m_dist_awgn;
        std::vector<uint8_t> data(c * k), cw(c * n);
        std::vector<float> ch(c * n);
        std::vector<int8_t> dec(c * n);
        for (size_t f = 0; f < c; f++) fill_with_random(data.data() + f * k, (unsigned)k);
        m_encoder->encode_batch(data.data(), c, cw.data());
        for (size_t f = 0; f < c; f++) make_frame(EbNo_dB, data.data() + f * k, cw.data() + f * n, ch.data() + f * n);
        m_decoder->decode_batch(ch.data(), c, dec.data(), nullptr, nullptr);
        size_t first = c;
        unsigned first_errors = 0;
        for (size_t f = 0; f < c; f++) {
            const unsigned e = compare_data(cw.data() + f * n, dec.data() + f * n);
            if (e) {
                first = f;
                first_errors = e;
                break;
            }
        }
        if (first == c) {
            m_num_bits += (uint64_t)c * k;
            done += c;
            chunk = std::min<size_t>(chunk * 2, 4096);
            continue;
        }
        // rewind and consume exactly frames 0..first of the chunk
        m_lfsr = lfsr;
        m_lfsr_index = lfsr_index;
        m_generator_awgn = gen;
        m_dist_awgn = dist;
        std::vector<uint8_t> d1(k), c1(n);
        std::vector<float> ch1(n);
        for (size_t f = 0; f <= first; f++) {
            fill_with_random(d1.data(), (unsigned)k);
            const uint8_t* cwf = cw.data() + f * n;   // same data word -> same codeword
            std::copy(cwf, cwf + n, c1.begin());
            make_frame(EbNo_dB, d1.data(), c1.data(), ch1.data());
        }
        m_num_bits += (uint64_t)(first + 1) * k;
        m_num_errors += first_errors;
        if (frames_run) *frames_run = done + first + 1;
        return true;
    }
    if (frames_run) *frames_run = done;
    return false;
}

}  // namespace wimax_ldpc_lib
