// decode_g0.hip -- mixed-batch group decode kernels (all rates of one z) for z = 24, 28, 32, 36.
#include "group.hpp"
WL_DEFINE_GROUP(24)
WL_DEFINE_GROUP(28)
WL_DEFINE_GROUP(32)
WL_DEFINE_GROUP(36)
