/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 13, 29, 45, 61 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(13, 16, 16)
CS_FAST64_INSTANTIATE(13, 32, 16)
CS_FAST64_INSTANTIATE(29, 16, 16)
CS_FAST64_INSTANTIATE(29, 32, 16)
CS_FAST64_INSTANTIATE(45, 16, 16)
CS_FAST64_INSTANTIATE(45, 32, 16)
CS_FAST64_INSTANTIATE(61, 16, 16)
CS_FAST64_INSTANTIATE(61, 32, 16)
