// Set-up, seam and FP32 apply kernels of degree 4 (Gauss points 5 per direction); see kernels.hpp.
#include "launch.hpp"

IGAK_DEFINE_DEGREE(4)
