// General rational (per-function weight) apply kernels of degree 2.
#include "launch.hpp"

IGAK_DEFINE_RAT(2)
