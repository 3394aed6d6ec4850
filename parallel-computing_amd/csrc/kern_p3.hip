// Set-up, seam and FP32 apply kernels of degree 3 (Gauss points 4 per direction); see kernels.hpp.
#include "launch.hpp"

IGAK_DEFINE_DEGREE(3)
