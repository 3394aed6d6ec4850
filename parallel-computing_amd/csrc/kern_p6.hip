// Set-up, seam and FP32 apply kernels of degree 6 (Gauss points 7 per direction); see kernels.hpp.
#include "launch.hpp"

IGAK_DEFINE_DEGREE(6)
