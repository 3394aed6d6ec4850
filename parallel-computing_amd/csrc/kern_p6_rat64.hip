// General rational (per-function weight) FP64 apply kernels of degree 6.
#include "launch.hpp"

IGAK_DEFINE_RAT64(6)
