// FP64 apply kernels of degree 3.
#include "launch.hpp"

IGAK_DEFINE_F64(3)
