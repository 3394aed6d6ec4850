// General rational (per-function weight) FP64 apply kernels of degree 3.
#include "launch.hpp"

IGAK_DEFINE_RAT64(3)
