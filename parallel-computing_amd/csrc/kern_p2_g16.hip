// FP32 apply kernels of degree 2 on FP16-stored geometry factors.
#include "launch.hpp"

IGAK_DEFINE_G16(2)
