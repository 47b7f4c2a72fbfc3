// decode_g1.hip -- mixed-batch group decode kernels (all rates of one z) for z = 40, 44, 48, 52.
#include "group.hpp"
WL_DEFINE_GROUP(40)
WL_DEFINE_GROUP(44)
WL_DEFINE_GROUP(48)
WL_DEFINE_GROUP(52)
