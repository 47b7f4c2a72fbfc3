// decode_g3.hip -- mixed-batch group decode kernels (all rates of one z) for z = 72, 76, 80, 84.
#include "group.hpp"
WL_DEFINE_GROUP(72)
WL_DEFINE_GROUP(76)
WL_DEFINE_GROUP(80)
WL_DEFINE_GROUP(84)
