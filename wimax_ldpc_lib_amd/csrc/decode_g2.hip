// decode_g2.hip -- mixed-batch group decode kernels (all rates of one z) for z = 56, 60, 64, 68.
#include "group.hpp"
WL_DEFINE_GROUP(56)
WL_DEFINE_GROUP(60)
WL_DEFINE_GROUP(64)
WL_DEFINE_GROUP(68)
