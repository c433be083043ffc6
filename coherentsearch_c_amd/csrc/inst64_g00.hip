/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 1, 17, 33, 49 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(1, 16, 16)
CS_FAST64_INSTANTIATE(1, 32, 16)
CS_FAST64_INSTANTIATE(17, 16, 16)
CS_FAST64_INSTANTIATE(17, 32, 16)
CS_FAST64_INSTANTIATE(33, 16, 16)
CS_FAST64_INSTANTIATE(33, 32, 16)
CS_FAST64_INSTANTIATE(49, 16, 16)
CS_FAST64_INSTANTIATE(49, 32, 16)
