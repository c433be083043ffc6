/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 7, 23, 39, 55 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(7, 16, 16)
CS_FAST64_INSTANTIATE(7, 32, 16)
CS_FAST64_INSTANTIATE(23, 16, 16)
CS_FAST64_INSTANTIATE(23, 32, 16)
CS_FAST64_INSTANTIATE(39, 16, 16)
CS_FAST64_INSTANTIATE(39, 32, 16)
CS_FAST64_INSTANTIATE(55, 16, 16)
CS_FAST64_INSTANTIATE(55, 32, 16)
