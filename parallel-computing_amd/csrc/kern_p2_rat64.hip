// General rational (per-function weight) FP64 apply kernels of degree 2.
#include "launch.hpp"

IGAK_DEFINE_RAT64(2)
