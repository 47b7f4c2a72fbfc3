/* ldpc_decoder.h -- TDMP decoder, drop-in for the reference's
 * include/ldpc_decoder.h:12-34 (lib/ldpc_decoder.cc).  Decoding runs on the
 * GPU through the C ABI in wimax_ldpc_amd.h; results are bit-identical to the
 * reference's (posterior in rx, 0/1 bytes in decoded, same return value). */
#ifndef LDPC_DECODER_H
#define LDPC_DECODER_H

#include <stddef.h>
#include <stdint.h>

#include "enums.h"
#include "ldpc.h"

struct wl_decoder;

namespace wimax_ldpc_lib {

    class ldpc_decoder : public ldpc
    {
        private:
            unsigned int m_max_iter;
            coderate     m_rate;
            wl_decoder*  m_handle;
            unsigned int m_flags;

            void recreate(unsigned int flags);

        public:
            ldpc_decoder();
            ldpc_decoder(coderate rate,
                         unsigned int z_factor,
                         unsigned int max_iter,
                         unsigned int num_threads);
            ~ldpc_decoder();

            unsigned int compute_syndrome(int8_t* rx_codeword, bool exit_early);
            unsigned int decode(int8_t* rx_codeword, int8_t* decoded);
            unsigned int decode(float* rx_codeword, int8_t* decoded);

            /* Extension: nframes frames in one GPU launch, with decode()'s
             * semantics per frame (rx [nframes][N] replaced by posteriors,
             * decoded [nframes][N] 0/1 bytes, ret [nframes] return values,
             * iters [nframes] iteration counts, nullable). */
            void decode_batch(float* rx, size_t nframes, int8_t* decoded, uint8_t* ret, uint8_t* iters);

            /* Extension: decode(int8_t* rx, int8_t* decoded) for nframes
             * frames of 0/1 bytes in one GPU launch (rx is not modified). */
            void decode_batch(const int8_t* rx, size_t nframes, int8_t* decoded, uint8_t* ret, uint8_t* iters);

            /* Extension: false = exactly max_iter iterations, no syndrome checks. */
            void set_early_stop(bool early_stop);

            /* Extension: the message rule the reference selects at compile
             * time (lib/ldpc_decoder.cc:17-21,215-238): RULE_NMS normalised
             * min-sum (the default, as shipped), RULE_OMS offset min-sum,
             * RULE_MS plain min-sum. */
            enum message_rule { RULE_NMS = 0, RULE_OMS = 1, RULE_MS = 2 };
            void set_message_rule(message_rule rule);

            wl_decoder* handle() { return m_handle; }
    };
}  // end namespace wimax_ldpc_lib

#endif  // LDPC_DECODER_H
