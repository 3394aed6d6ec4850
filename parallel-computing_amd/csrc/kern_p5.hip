// Set-up, seam and FP32 apply kernels of degree 5 (Gauss points 6 per direction); see kernels.hpp.
#include "launch.hpp"

IGAK_DEFINE_DEGREE(5)
