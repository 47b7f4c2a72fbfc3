// decode_r4_lo.hip -- TDMP decoder kernels for coderate 4 at lifting factors 24..60.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE_lo(4)
