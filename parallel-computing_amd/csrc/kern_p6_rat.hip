// General rational (per-function weight) apply kernels of degree 6.
#include "launch.hpp"

IGAK_DEFINE_RAT(6)
