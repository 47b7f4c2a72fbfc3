// decode_p5.hip -- small-batch TDMP decoder kernels (two lanes per row) for coderate 5 at all 19 lifting factors.
#include "tdmp_pair.hpp"
WL_DEFINE_PAIR_TABLE(5)
