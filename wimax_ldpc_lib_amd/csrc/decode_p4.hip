// decode_p4.hip -- small-batch TDMP decoder kernels (two lanes per row) for coderate 4 at all 19 lifting factors.
#include "tdmp_pair.hpp"
WL_DEFINE_PAIR_TABLE(4)
