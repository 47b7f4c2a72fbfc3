/* ldpc.h -- code object, drop-in for the reference's include/ldpc.h:35-61.
 *
 * Same class, constructor and members.  The expanded parity-check table
 * (m_checknode_array, [m_row_size][m_col_size], -1 padded) is built at
 * construction from the 802.16e base matrices instead of being linked in as
 * the 114 generated arrays of the reference's include/matrices.h. */
#ifndef LDPC_H
#define LDPC_H

#include <stdint.h>
#include <stdio.h>

#include <chrono>
#include <iostream>
#include <vector>

#include "enums.h"

#define BASE_LDPC_BLOCK_LEN 2304

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

            std::vector<int16_t> m_table;

        public:
            ldpc();
            /* lib/ldpc.cc:11-63: throws std::exception on an invalid rate or z */
            ldpc(coderate rate, unsigned int z_factor, unsigned int num_threads);
            ~ldpc();

            uint64_t get_nanoseconds();     /* lib/ldpc.cc:1208-1215 */
    };
}  // end namespace wimax_ldpc_lib

#endif  // LDPC_H
