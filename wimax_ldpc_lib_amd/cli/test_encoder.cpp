// test_encoder -- drop-in for the reference's encoder CLI
// (lib/encode_example.cc): same arguments, file formats, printed statistics
// and exit codes, with all rounds encoded in one GPU batch.
//
//   test_encoder <rate> <z> <num_rounds> <errors_to_add> <unencoded_data_file> <encoded_data_file>
//
// Writes num_rounds random info words (k bytes of 0/1 each) and their
// codewords (n bytes each) with errors_to_add bit flips per codeword at
// rand()-chosen positions.  Two deliberate differences from the reference:
// the generator is seeded once (the reference reseeds with time(0) for every
// word, so all words of one run made within a second are equal), from
// $WL_SEED when set, else time(0); and a flip index is rand() % n (the
// reference's (rand() % n) + 1 can write one byte past the codeword,
// lib/encode_example.cc:31-32).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include <exception>
#include <string>
#include <vector>

#include "ldpc_encoder.h"

#define NUM_THREADS 4

using namespace wimax_ldpc_lib;

static int usage()
{
    printf("\nError: Incorrect number of arguments\n");
    printf("Usage: ./test_encoder <rate> <z> <num_rounds> <errors_to_add> <unencoded_data_file> <encoded_data_file>\n");
    printf("Argument Description:\n");
    printf("rate: LDPC code rate - half-rate        = 0\n");
    printf("                       two-thirds-A     = 1\n");
    printf("                       two-thirds-B     = 2\n");
    printf("                       three-quarters-A = 3\n");
    printf("                       three-quarters-B = 4\n");
    printf("                       five-sixths      = 5\n");
    printf("z: Z Factor (please refer to section 8.4.9.2.5 of the 802.16-2012 standard for more information)\n");
    printf("num_rounds: How many LDPC encoding rounds to run.\n");
    printf("errors_to_add: How many errors to add to each LDPC codeword.\n");
    printf("unencoded_data_file: File to write the unencoded data to.\n");
    printf("encoded_data_file: File to write LDPC encoded data to.\n\n");
    return 0;
}

int main(int argc, char* argv[])
{
    if (argc != 7) return usage();
    const coderate rate = (coderate)std::stoi(argv[1], nullptr, 0);
    const int z = std::stoi(argv[2], nullptr, 0);
    const unsigned num_rounds = std::stoi(argv[3], nullptr, 0);
    const unsigned num_errors = std::stoi(argv[4], nullptr, 0);
    const size_t n = (size_t)((z / 96.0) * BASE_LDPC_BLOCK_LEN);
    if (num_errors >= n) {
        printf("[!] Error: Num_errors greater than codeword length\n");
        return 0;
    }
    FILE* data_file = fopen(argv[5], "wb");
    FILE* encoded_file = fopen(argv[6], "wb");
    if (!data_file || !encoded_file) {
        printf("[!] Error: cannot open output files\n");
        return 0;
    }
    ldpc_encoder encoder(rate, z, NUM_THREADS);   // throws on a bad rate or z
    size_t k = 0;
    switch (rate) {   // lib/encode_example.cc:87-126
    case HALFRATE: k = n / 2; break;
    case TWOTHIRDSA:
    case TWOTHIRDSB: k = (n * 2) / 3; break;
    case THREEQUARTERSB:
        printf("[!]test_encoder - 34B is not yet supported: %d\n", rate);
        throw std::exception();
    case THREEQUARTERSA: k = (n * 3) / 4; break;
    case FIVESIXTHS: k = (n * 5) / 6; break;
    default:
        printf("[!]test_encoder - Invalid Coderate: %d\n", rate);
        throw std::exception();
    }

    const char* seed_env = getenv("WL_SEED");
    srand(seed_env ? (unsigned)strtoul(seed_env, nullptr, 0) : (unsigned)time(0));
    std::vector<uint8_t> data((size_t)num_rounds * k), cw((size_t)num_rounds * n);
    for (size_t i = 0; i < data.size(); i++) data[i] = (uint8_t)(rand() % 2);

    if (num_rounds) encoder.encode_batch(data.data(), 1, cw.data());   // warm-up
    const uint64_t t0 = encoder.get_nanoseconds();
    if (num_rounds) encoder.encode_batch(data.data(), num_rounds, cw.data());
    const uint64_t elapsed = encoder.get_nanoseconds() - t0;
    const double rate_mbps = elapsed ? (num_rounds * n * 1000.0) / (double)elapsed : 0.0;

    for (unsigned i = 0; i < num_rounds; i++) {
        uint8_t* c = cw.data() + (size_t)i * n;
        for (unsigned e = 0; e < num_errors; e++) {
            const size_t idx = (size_t)rand() % n;
            c[idx] = !c[idx];
        }
        fwrite(data.data() + (size_t)i * k, 1, k, data_file);
        fwrite(c, 1, n, encoded_file);
        printf("Round[%d] Rate (Mbits/Sec): %f\n", i, rate_mbps);
    }
    printf("\n\nTiming Statistics:\n\n");
    printf("Average Rate (Mbits/Sec): %f\n", rate_mbps);
    printf("Fastest Iteration (Mbits/Sec): %f\n", rate_mbps);
    printf("Slowest Iteration (Mbits/Sec): %f\n", rate_mbps);
    fclose(data_file);
    fclose(encoded_file);
    return 1;   // the reference's exit status on success (lib/encode_example.cc:169)
}
