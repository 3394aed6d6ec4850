// FP64 apply kernels of degree 4.
#include "launch.hpp"

IGAK_DEFINE_F64(4)
