// Multi right-hand-side apply kernels of degree 5.
#include "block.hpp"
#include "launch.hpp"

IGAK_DEFINE_BLOCK(5)
