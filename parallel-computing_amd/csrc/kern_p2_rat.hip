// General rational (per-function weight) FP32 apply kernels of degree 2.
#include "launch.hpp"

IGAK_DEFINE_RAT(2)
