// decode_r1.hip -- TDMP decoder kernels for coderate 1 at all 19 lifting factors.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE(1)
