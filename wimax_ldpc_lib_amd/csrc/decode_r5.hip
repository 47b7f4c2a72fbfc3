// decode_r5.hip -- TDMP decoder kernels for coderate 5 at all 19 lifting factors.
#include "dispatch.hpp"
WL_DEFINE_DECODE_TABLE(5)
