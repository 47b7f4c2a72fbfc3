// decode_r2.hip -- TDMP decoder kernels for coderate 2 at all 19 lifting factors.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE(2)
