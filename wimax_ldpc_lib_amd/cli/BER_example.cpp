// BER_example -- drop-in for the reference's BER sweep (lib/BER_example.cc):
// same arguments, the same BER_runner frame stream, the same per-point
// sample sizes and printed lines; decoding runs on the GPU in batches, and
// the curve is written as CSV (BER_LDPC_RATE<r>_N<n>_K<k>.csv: ebno_db,ber)
// instead of a MathGL plot.
//
//   BER_example <num_threads> <rate> <z> <EbNo_start> <EbNo_end> <EbNo_step>
//
// As in the reference, EbNo_start and EbNo_end are read as integers
// (lib/BER_example.cc:43-44) and each point first runs 46080 / n frames, then
// enough frames for 10 x 3.912 / BER bits (3.912e8 bits when no error was
// seen), continuing one frame stream across points.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <exception>
#include <string>
#include <vector>

#include "BER_runner.h"

#define MAX_ITER 10

using namespace wimax_ldpc_lib;

// nframes consecutive frames of the runner's stream, in GPU batches
static double run_frames(BER_runner& r, double ebno, unsigned long long nframes)
{
    double ber = 0.0;
    const unsigned long long chunk = 65536;
    for (unsigned long long done = 0; done < nframes; done += chunk)
        ber = r.run_batch(ebno, (size_t)std::min(chunk, nframes - done));
    if (nframes == 0) ber = r.run_batch(ebno, 0);
    return ber;
}

int main(int argc, char* argv[])
{
    if (argc != 7) {
        printf("\nError: Incorrect number of arguments\n");
        printf("Usage: ./BER_example <num_threads> <rate> <z> <EbNo_start> <EbNo_end> <EbNo_step> \n");
        printf("Argument Description:\n");
        printf("num_threads: Number of threads to use.  num_codewords modulo num_threads == 0.\n");
        printf("rate: LDPC code rate - half-rate        = 0\n");
        printf("                       two-thirds-A     = 1\n");
        printf("                       two-thirds-B     = 2\n");
        printf("                       three-quarters-A = 3\n");
        printf("                       three-quarters-B = 4\n");
        printf("                       five-sixths      = 5\n");
        printf("z: Z Factor (please refer to section 8.4.9.2.5 of the 802.16-2012 standard for more information)\n");
        printf("EbNo_start: Starting value for BER plot x axis (dB)\n");
        printf("EbNo_end: Ending value for BER plot x axis (dB)\n");
        printf("EbNo_step: Step value for BER plot x axis (dB)\n\n");
        return 0;
    }
    const coderate rate = (coderate)std::stoi(argv[2], nullptr, 0);
    const unsigned num_threads = std::stoi(argv[1], nullptr, 0);
    const int z = std::stoi(argv[3], nullptr, 0);
    const int start = std::stoi(argv[4], nullptr, 0);
    const int stop = std::stoi(argv[5], nullptr, 0);
    const float step = std::stof(argv[6], nullptr);

    BER_runner ber_test(rate, z, num_threads, MAX_ITER);
    const unsigned n = ber_test.get_codeword_len(), k = ber_test.get_dataword_len();

    unsigned num_steps = 0;
    if (fmod(stop - start, step) != 0)
        printf("[!] BER_example: step size does not evenly divide range\n");
    else
        num_steps = (unsigned)((stop - start) / step) + 1;

    char name[96];
    snprintf(name, sizeof name, "BER_LDPC_RATE%d_N%u_K%u.csv", (int)rate, n, k);
    FILE* csv = fopen(name, "w");
    if (!csv) {
        printf("[!] BER_example: cannot write %s\n", name);
        return 0;
    }
    fprintf(csv, "ebno_db,ber,bits,errors\n");

    double EbNo_dB = (double)start;
    for (unsigned i = 0; i < num_steps; i++) {
        printf("EbN0(dB): %f\n", EbNo_dB);
        double BER = run_frames(ber_test, EbNo_dB, 46080 / n);
        unsigned num_req;
        if (BER == 0.0)
            num_req = (unsigned)floor(3.912 / 0.00000001);
        else
            num_req = (unsigned)(10 * floor(3.912 / BER));
        BER = run_frames(ber_test, EbNo_dB, num_req / n);
        fprintf(csv, "%.6f,%.9e,%llu,%llu\n", EbNo_dB, BER, (unsigned long long)ber_test.get_total_num_bits(),
                (unsigned long long)ber_test.get_total_num_errors());
        ber_test.reset_test();
        printf("BER: %e\n\n", BER);
        EbNo_dB += step;
    }
    fclose(csv);
    return 1;   // the reference's exit status (lib/BER_example.cc:187)
}
