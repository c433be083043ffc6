/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 8, 24, 40, 56 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(8, 16, 16)
CS_FAST64_INSTANTIATE(8, 32, 16)
CS_FAST64_INSTANTIATE(24, 16, 16)
CS_FAST64_INSTANTIATE(24, 32, 16)
CS_FAST64_INSTANTIATE(40, 16, 16)
CS_FAST64_INSTANTIATE(40, 32, 16)
CS_FAST64_INSTANTIATE(56, 16, 16)
CS_FAST64_INSTANTIATE(56, 32, 16)
