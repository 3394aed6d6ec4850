// General rational (per-function weight) apply kernels of degree 3.
#include "launch.hpp"

IGAK_DEFINE_RAT(3)
