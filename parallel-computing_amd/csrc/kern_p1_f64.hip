// FP64 apply kernels of degree 1.
#include "launch.hpp"

IGAK_DEFINE_F64(1)
