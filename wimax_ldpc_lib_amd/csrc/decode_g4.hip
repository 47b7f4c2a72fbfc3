// decode_g4.hip -- mixed-batch group decode kernels (all rates of one z) for z = 88, 92, 96.
#include "group.hpp"
WL_DEFINE_GROUP(88)
WL_DEFINE_GROUP(92)
WL_DEFINE_GROUP(96)
