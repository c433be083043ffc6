/* Fused search kernel instantiations for nharms = 60 (m = 16, 32; k <= 6). */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(60, 16, 6)
CS_FAST_INSTANTIATE(60, 32, 6)
