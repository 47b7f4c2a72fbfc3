/* BER_runner.h -- bit-error-rate harness, drop-in for the reference's
 * include/BER_runner.h:11-69 (lib/BER_runner.cc).
 *
 * run_iteration() keeps the reference's frame source and channel exactly:
 * the x^31 + x^28 + 1 LFSR (seed 0xdeadbeef, reset every 115200 bits), the
 * encoder, bit b -> 2b-1, std::mt19937(0) with std::normal_distribution<float>
 * noise of amplitude sqrt(0.5)/sqrt(EbN0)/sqrt(R), decoding of the raw
 * channel values, and errors counted over the n coded bits against k info
 * bits per frame.  With the same C++ standard library the frames are the
 * reference's frames, and the GPU decoder returns the reference's decisions.
 * run_batch() is the same stream of frames, nframes at a time, decoded in
 * one GPU launch. */
#ifndef BER_RUNNER_H
#define BER_RUNNER_H

#include <stddef.h>
#include <stdint.h>

#include <random>
#include <vector>

#include "enums.h"
#include "ldpc.h"
#include "ldpc_decoder.h"
#include "ldpc_encoder.h"

namespace wimax_ldpc_lib {
    class BER_runner : public ldpc
    {
        private:
            ldpc_encoder* m_encoder;
            ldpc_decoder* m_decoder;

            uint64_t m_num_bits;
            uint64_t m_num_errors;

            unsigned int m_codeword_len;
            unsigned int m_dataword_len;

            float m_float_rate;

            uint32_t m_lfsr;
            uint32_t m_lfsr_start_state;
            unsigned int m_lfsr_index;
            unsigned int m_lfsr_reset_len;

            std::mt19937 m_generator_awgn;
            std::normal_distribution<float> m_dist_awgn;

            uint32_t lfsr_get_num();
            void fill_with_random(uint8_t* buffer, unsigned int buf_len);
            unsigned int compare_data(const uint8_t* original_data, const int8_t* decoded_data);
            void make_frame(double EbNo_dB, uint8_t* dataword, uint8_t* codeword, float* channel);
            void require_code() const;

        public:
            /* include/BER_runner.h:51: no code; run_iteration / run_batch /
             * run_until_error throw std::exception on such an object */
            BER_runner();
            BER_runner(coderate rate, unsigned int z_factor,
                       unsigned int num_threads,
                       unsigned int max_iter);
            ~BER_runner();

            void reset_test();
            double run_iteration(double EbNo_dB);
            /* Extension: nframes consecutive run_iteration() frames in one
             * GPU decode; returns the running BER. */
            double run_batch(double EbNo_dB, size_t nframes);

            /* Extension: the loop "run_iteration() up to max_frames times,
             * stop after the first frame with a bit error"
             * (lib/Capacity_plot_example.cc:147-154), decoded in GPU batches.
             * Leaves the frame source and the counters exactly where that
             * loop would; returns true if an error occurred, and the number
             * of frames the loop ran in *frames_run (nullable). */
            bool run_until_error(double EbNo_dB, size_t max_frames, size_t* frames_run);

            uint64_t get_total_num_bits();
            uint64_t get_total_num_errors();
            unsigned int get_codeword_len();
            unsigned int get_dataword_len();
    };
}  // end namespace wimax_ldpc_lib

#endif  // BER_RUNNER_H
