// decode_p0.hip -- small-batch TDMP decoder kernels (two lanes per row) for coderate 0 at all 19 lifting factors.
#include "tdmp_pair.hpp"
WL_DEFINE_PAIR_TABLE(0)
