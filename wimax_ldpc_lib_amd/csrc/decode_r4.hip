// decode_r4.hip -- TDMP decoder kernels for coderate 4 at all 19 lifting factors.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE(4)
