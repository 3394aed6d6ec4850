// General rational (per-function weight) FP32 apply kernels of degree 4.
#include "launch.hpp"

IGAK_DEFINE_RAT(4)
