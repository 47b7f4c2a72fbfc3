// matrices.hpp -- lookup of the exported tables of include/matrices.h.
#pragma once

#include <stdint.h>

namespace wl {

// &wimax_<n>_<rate>[0][0] of code (rate, z), nullptr for an invalid code
int16_t* matrix_table(int rate, int z);

}  // namespace wl
