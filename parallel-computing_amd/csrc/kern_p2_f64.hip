// FP64 apply kernels of degree 2.
#include "launch.hpp"

IGAK_DEFINE_F64(2)
