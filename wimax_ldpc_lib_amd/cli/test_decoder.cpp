// test_decoder -- drop-in for the reference's file decoder CLI
// (lib/decode_example.cc): same arguments, file formats, printed statistics
// and exit codes, with every frame of the file decoded in one GPU batch.
//
//   test_decoder <num_threads> <rate> <z> <num_codewords> <encoded_data_file> <decoded_data_file>
//
// The input file holds num_codewords frames of n bytes, one 0/1 byte per
// code bit (what test_encoder writes); each is decoded through the
// hard-input path (soft = 2 bit - 1, lib/ldpc_decoder.cc:108-118) with at
// most 10 iterations and the syndrome early stop, and the first k bytes of
// each decoded frame are written out.  As in the reference, frames are taken
// in groups of num_threads and a trailing partial group is not decoded
// (lib/decode_example.cc:118); num_threads only sets that grouping and the
// per-group report, the decoding itself is one launch.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <exception>
#include <string>
#include <vector>

#include "ldpc_decoder.h"

#define MAX_ITER 10

using namespace wimax_ldpc_lib;

static int usage()
{
    printf("\nError: Incorrect number of arguments\n");
    printf("Usage: ./test_decoder <num_threads> <rate> <z> <num_codewords> <encoded_data_file> <decoded_data_file>\n");
    printf("Argument Description:\n");
    printf("num_threads: Frames per reported group.  num_codewords modulo num_threads == 0.\n");
    printf("rate: LDPC code rate - half-rate        = 0\n");
    printf("                       two-thirds-A     = 1\n");
    printf("                       two-thirds-B     = 2\n");
    printf("                       three-quarters-A = 3\n");
    printf("                       three-quarters-B = 4\n");
    printf("                       five-sixths      = 5\n");
    printf("z: Z Factor (please refer to section 8.4.9.2.5 of the 802.16-2012 standard for more information)\n");
    printf("num_codewords: How many LDPC codewords are in the file.\n");
    printf("encoded_data_file: File to read the encoded data from.\n");
    printf("decoded_data_file: File to write decoded data to.\n\n");
    return 0;
}

int main(int argc, char* argv[])
{
    if (argc != 7) return usage();
    const unsigned num_threads = std::stoi(argv[1], nullptr, 0);
    const coderate rate = (coderate)std::stoi(argv[2], nullptr, 0);
    const int z = std::stoi(argv[3], nullptr, 0);
    const unsigned num_codewords = std::stoi(argv[4], nullptr, 0);
    if (num_threads == 0) return usage();

    const size_t n = (size_t)((z / 96.0) * BASE_LDPC_BLOCK_LEN);
    std::vector<int8_t> file_buffer((size_t)num_codewords * n);
    FILE* encoded_data = fopen(argv[5], "rb");
    FILE* decoded_data = fopen(argv[6], "wb");
    if (encoded_data) {
        const size_t got = fread(file_buffer.data(), 1, file_buffer.size(), encoded_data);
        fclose(encoded_data);
        if (got != file_buffer.size()) {
            printf("[!] Error: File does not contain enough data.\n");
            return 0;
        }
    }
    if (!decoded_data) {
        printf("[!] Error: cannot open %s\n", argv[6]);
        return 0;
    }

    // the reference reads the file before it builds its decoders (lib/decode_example.cc:60-76)
    ldpc_decoder dec(rate, z, MAX_ITER, num_threads);   // throws on a bad rate or z, as the reference
    size_t k = 0;
    switch (rate) {   // lib/decode_example.cc:81-115
    case HALFRATE: k = n / 2; break;
    case TWOTHIRDSA:
    case TWOTHIRDSB: k = (n * 2) / 3; break;
    case THREEQUARTERSA:
    case THREEQUARTERSB: k = (n * 3) / 4; break;
    case FIVESIXTHS: k = (n * 5) / 6; break;
    default:
        printf("[!]test_encoder - Invalid Coderate: %d\n", rate);
        throw std::exception();
    }

    const size_t groups = num_codewords / num_threads;
    const size_t frames = groups * num_threads;
    std::vector<unsigned> initial(frames);
    for (size_t f = 0; f < frames; f++) initial[f] = dec.compute_syndrome(file_buffer.data() + f * n, false);

    std::vector<int8_t> decoded(frames * n);
    std::vector<uint8_t> ret(frames);
    if (frames) dec.decode_batch(file_buffer.data(), frames, decoded.data(), ret.data(), nullptr);   // warm-up
    const uint64_t t0 = dec.get_nanoseconds();
    if (frames) dec.decode_batch(file_buffer.data(), frames, decoded.data(), ret.data(), nullptr);
    const uint64_t elapsed = dec.get_nanoseconds() - t0;

    // one timing for the whole batch: each group is reported at the batch rate
    const double batch_rate = elapsed ? (frames * n * 1000.0) / (double)elapsed : 0.0;
    for (size_t g = 0; g < groups; g++) {
        for (unsigned j = 0; j < num_threads; j++)
            fwrite(decoded.data() + (g * num_threads + j) * n, 1, k, decoded_data);
        printf("Rate (Mbits/Sec): %f\n", batch_rate);
        for (unsigned j = 0; j < num_threads; j++) {
            printf("Initial Number of Errors (%u): %u\n", j, initial[g * num_threads + j]);
            printf("Final Number of Errors (%u): %u\n", j, (unsigned)ret[g * num_threads + j]);
        }
        printf("\n");
    }
    printf("\nTiming Statistics:\n\n");
    printf("Average Rate (Mbits/Sec): %f\n", batch_rate);
    printf("Fastest Iteration (Mbits/Sec): %f\n", batch_rate);
    printf("Slowest Iteration (Mbits/Sec): %f\n", batch_rate);
    fclose(decoded_data);
    return 1;   // the reference's exit status on success (lib/decode_example.cc:178)
}
