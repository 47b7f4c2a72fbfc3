// decode_r2_lo.hip -- TDMP decoder kernels for coderate 2 at lifting factors 24..60.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE_lo(2)
