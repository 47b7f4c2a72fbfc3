// qc_host.hpp -- host-side description of a quasi-cyclic LDPC code given at
// run time (base matrix of circulant shifts, or an alist file).
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/wimax_ldpc_amd.h"
#include "generic.hpp"

struct wl_qc_code {
    int nb = 0, mb = 0, z = 0;
    std::vector<int16_t> base;    // [mb][nb] circulant shifts, -1 = zero block
    wl_code_info info{};          // rate -1; d = x = -1 without the dual-diagonal parity part
    std::vector<int16_t> table;   // [m][c] variables of each check row, ascending, -1 padded
    wl::GenDesc desc{};           // the decoder's descriptor (host copy)
};

namespace wl {

// Validates the base matrix, fills info, table and desc, and detects the
// 802.16e / 802.11n dual-diagonal parity structure (sets info.x, info.d).
int qc_build(int nb, int mb, int z, const int16_t* base, wl_qc_code& q);

}  // namespace wl
