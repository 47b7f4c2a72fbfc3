// decode_r0.hip -- TDMP decoder kernels for coderate 0 at all 19 lifting factors.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE(0)
