// General rational (per-function weight) FP64 apply kernels of degree 4.
#include "launch.hpp"

IGAK_DEFINE_RAT64(4)
