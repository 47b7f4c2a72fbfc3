// decode_r3.hip -- TDMP decoder kernels for coderate 3 at all 19 lifting factors.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE(3)
