// Set-up, seam and FP32 apply kernels of degree 1 (Gauss points 2 per direction); see kernels.hpp.
#include "launch.hpp"

IGAK_DEFINE_DEGREE(1)
