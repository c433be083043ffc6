/* Fused search kernel instantiations for nharms = 32 (m = 16, 32; k <= 6). */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(32, 16, 6)
CS_FAST_INSTANTIATE(32, 32, 6)
