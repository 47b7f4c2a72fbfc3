/* ldpc.h -- code object, drop-in for the reference's include/ldpc.h:35-61.
 *
 * Same class, constructor and members.  m_checknode_array points at the
 * code's expanded parity-check table of matrices.h ([m_row_size][m_col_size],
 * -1 padded), as the reference's set_checknode_array does
 * (lib/ldpc.cc:69-1206); the library fills those tables from the 802.16e base
 * matrices when it is loaded. */
#ifndef LDPC_H
#define LDPC_H

#include <stdint.h>
#include <stdio.h>

#include <chrono>
#include <iostream>

#include "enums.h"
#include "matrices.h"

#define BASE_LDPC_BLOCK_LEN 2304

/* lifting sizes compiled in (include/ldpc.h:14-32 of the reference) */
#define Z_24
#define Z_28
#define Z_32
#define Z_36
#define Z_40
#define Z_44
#define Z_48
#define Z_52
#define Z_56
#define Z_60
#define Z_64
#define Z_68
#define Z_72
#define Z_76
#define Z_80
#define Z_84
#define Z_88
#define Z_92
#define Z_96

namespace wimax_ldpc_lib {
    class ldpc
    {
        protected:
            int16_t*     m_checknode_array;
            unsigned int m_D;
            unsigned int m_N;
            unsigned int m_M;
            unsigned int m_X;
            unsigned int m_z;
            unsigned int m_col_size;
            unsigned int m_row_size;
            unsigned int m_num_threads;

        public:
            ldpc();
            /* lib/ldpc.cc:11-63: throws std::exception on an invalid rate or z */
            ldpc(coderate rate, unsigned int z_factor, unsigned int num_threads);
            ~ldpc();

            uint64_t get_nanoseconds();     /* lib/ldpc.cc:1208-1215 */
    };
}  // end namespace wimax_ldpc_lib

#endif  // LDPC_H
