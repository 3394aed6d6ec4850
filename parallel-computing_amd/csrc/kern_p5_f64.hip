// FP64 apply kernels of degree 5.
#include "launch.hpp"

IGAK_DEFINE_F64(5)
