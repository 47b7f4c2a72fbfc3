// decode_r3_lo.hip -- TDMP decoder kernels for coderate 3 at lifting factors 24..60.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE_lo(3)
