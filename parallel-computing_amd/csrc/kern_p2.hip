// Set-up, seam and FP32 apply kernels of degree 2 (Gauss points 3 per direction); see kernels.hpp.
#include "launch.hpp"

IGAK_DEFINE_DEGREE(2)
