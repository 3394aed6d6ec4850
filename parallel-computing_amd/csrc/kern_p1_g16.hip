// FP32 apply kernels of degree 1 on FP16-stored geometry factors.
#include "launch.hpp"

IGAK_DEFINE_G16(1)
