// FP64 apply kernels of degree 6.
#include "launch.hpp"

IGAK_DEFINE_F64(6)
