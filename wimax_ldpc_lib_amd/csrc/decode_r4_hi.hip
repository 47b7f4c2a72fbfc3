// decode_r4_hi.hip -- TDMP decoder kernels for coderate 4 at lifting factors 64..96.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE_hi(4)
