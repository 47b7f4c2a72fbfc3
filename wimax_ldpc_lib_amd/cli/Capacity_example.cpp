// Capacity_example -- drop-in for the reference's Eb/N0 threshold search
// (lib/Capacity_plot_example.cc): for the rates 1/2, 2/3A, 2/3B, 3/4A and 5/6
// at one z, the smallest Eb/N0 on a 0.01 dB grid from 0.25 dB at which
// 3.912e5 / n + 1 consecutive frames of the BER_runner stream decode without
// a bit error (BER 1e-5 at 98 % confidence).  Same arguments, search, frame
// stream and printed lines; decoding runs on the GPU in batches
// (BER_runner::run_until_error), and the points are written as CSV
// (CAP_LDPC_N<n>.csv) with the Shannon and BPSK-AWGN limits at each rate
// instead of a MathGL plot.
//
//   Capacity_example <num_threads> <z>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <string>

#include "BER_runner.h"

#define MAX_ITER 10
#define EBNO_STEP 0.01

using namespace wimax_ldpc_lib;

// capacity of BPSK on the real AWGN channel with noise variance s2, bits per
// dimension: 1 - E[log2(1 + exp(-2 y / s2))], y ~ N(1, s2) (trapezoid rule
// over +-12 standard deviations)
static double bpsk_capacity(double s2)
{
    const double sd = sqrt(s2);
    const int npts = 4000;
    const double lo = 1.0 - 12 * sd, hi = 1.0 + 12 * sd, h = (hi - lo) / npts;
    double acc = 0.0;
    for (int i = 0; i <= npts; i++) {
        const double y = lo + i * h;
        const double pdf = exp(-(y - 1.0) * (y - 1.0) / (2 * s2)) / sqrt(2 * M_PI * s2);
        const double t = -2.0 * y / s2;
        const double l = t > 30 ? t / log(2.0) : log2(1.0 + exp(t));
        acc += (i == 0 || i == npts ? 0.5 : 1.0) * pdf * l;
    }
    return 1.0 - acc * h;
}

// Eb/N0 (dB) at which BPSK-AWGN capacity equals R (bisection on sigma^2)
static double bpsk_limit_db(double R)
{
    double lo = 1e-3, hi = 1e3;   // noise variance
    for (int it = 0; it < 200; it++) {
        const double mid = sqrt(lo * hi);
        if (bpsk_capacity(mid) > R) lo = mid; else hi = mid;
    }
    const double s2 = sqrt(lo * hi);
    return 10 * log10(1.0 / (2 * R * s2));
}

// Shannon limit of the real AWGN channel at R bits per dimension
static double shannon_limit_db(double R)
{
    return 10 * log10((pow(2.0, 2 * R) - 1.0) / (2 * R));
}

int main(int argc, char* argv[])
{
    if (argc != 3) {
        printf("\nError: Incorrect number of arguments\n");
        printf("Usage: ./Capacity_example <num_threads> <z>\n");
        printf("Argument Description:\n");
        printf("num_threads: Number of threads to use.\n");
        printf("z: Z Factor (please refer to section 8.4.9.2.5 of the 802.16-2012 standard for more information)\n");
        return 0;
    }
    const unsigned num_threads = std::stoi(argv[1], nullptr, 0);
    const int z = std::stoi(argv[2], nullptr, 0);
    const unsigned ratelist[5] = {0, 1, 2, 3, 5};
    const char* names[6] = {"1/2", "2/3A", "2/3B", "3/4A", "3/4B", "5/6"};
    const double plot_rate[6] = {0.5, 0.67, 0.67, 0.75, 0.75, 0.83};   // lib/Capacity_plot_example.cc:109-140
    const unsigned num_req = (unsigned)(3.912 / 0.00001);

    char name[64];
    snprintf(name, sizeof name, "CAP_LDPC_N%u.csv", (unsigned)((z / 96.0) * 2304));
    FILE* csv = fopen(name, "w");
    if (!csv) {
        printf("[!] Capacity_example: cannot write %s\n", name);
        return 0;
    }
    fprintf(csv, "rate,plot_rate,ebno_db,shannon_limit_db,bpsk_limit_db,frames_run\n");
    for (unsigned i = 0; i < 5; i++) {
        const coderate rate = (coderate)ratelist[i];
        BER_runner ber_test(rate, z, num_threads, MAX_ITER);
        double EbNo_dB = 0.25;
        unsigned long long frames_total = 0;
        for (;;) {
            size_t ran = 0;
            const bool error = ber_test.run_until_error(EbNo_dB, num_req / ber_test.get_codeword_len() + 1, &ran);
            frames_total += ran;
            ber_test.reset_test();
            if (!error) {
                printf("Rate: %s, EbNo: %f\n", names[rate], EbNo_dB);
                break;
            }
            EbNo_dB += EBNO_STEP;
        }
        const double R = (double)ber_test.get_dataword_len() / ber_test.get_codeword_len();
        fprintf(csv, "%s,%.2f,%.6f,%.4f,%.4f,%llu\n", names[rate], plot_rate[rate], EbNo_dB, shannon_limit_db(R),
                bpsk_limit_db(R), frames_total);
    }
    fclose(csv);
    return 1;   // the reference's exit status (lib/Capacity_plot_example.cc:186)
}
