// General rational (per-function weight) FP32 apply kernels of degree 6.
#include "launch.hpp"

IGAK_DEFINE_RAT(6)
