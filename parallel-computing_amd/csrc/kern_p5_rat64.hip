// General rational (per-function weight) FP64 apply kernels of degree 5.
#include "launch.hpp"

IGAK_DEFINE_RAT64(5)
