/* ldpc_encoder.h -- dual-diagonal encoder, drop-in for the reference's
 * include/ldpc_encoder.h:10-27 (lib/ldpc_encoder.cc).  Encoding runs on the
 * GPU; codewords are bit-identical to the reference's. */
#ifndef LDPC_ENCODER_H
#define LDPC_ENCODER_H

#include <stddef.h>
#include <stdint.h>

#include "enums.h"
#include "ldpc.h"

struct wl_encoder;

namespace wimax_ldpc_lib {
    class ldpc_encoder : public ldpc
    {
        private:
            wl_encoder* m_handle;

        public:
            ldpc_encoder(coderate rate, unsigned int z_factor, unsigned int num_threads);
            /* Extension: standard = true selects the back-substitution
             * encoder that is valid for 3/4B as well (wl_encoder_create_standard);
             * for the other rates the codewords are the same. */
            ldpc_encoder(coderate rate, unsigned int z_factor, unsigned int num_threads, bool standard);
            ~ldpc_encoder();

            void encode_data(uint8_t* infoword, uint8_t* codeword);

            /* Extension: nframes frames, info [nframes][N-M], codeword [nframes][N]. */
            void encode_batch(const uint8_t* infoword, size_t nframes, uint8_t* codeword);

            wl_encoder* handle() { return m_handle; }
    };
}  // end namespace wimax_ldpc_lib

#endif  // LDPC_ENCODER_H
