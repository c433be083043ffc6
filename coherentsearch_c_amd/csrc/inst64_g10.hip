/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 11, 27, 43, 59 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(11, 16, 16)
CS_FAST64_INSTANTIATE(11, 32, 16)
CS_FAST64_INSTANTIATE(27, 16, 16)
CS_FAST64_INSTANTIATE(27, 32, 16)
CS_FAST64_INSTANTIATE(43, 16, 16)
CS_FAST64_INSTANTIATE(43, 32, 16)
CS_FAST64_INSTANTIATE(59, 16, 16)
CS_FAST64_INSTANTIATE(59, 32, 16)
