// General rational (per-function weight) FP32 apply kernels of degree 3.
#include "launch.hpp"

IGAK_DEFINE_RAT(3)
