/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 5, 21, 37, 53 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(5, 16, 16)
CS_FAST64_INSTANTIATE(5, 32, 16)
CS_FAST64_INSTANTIATE(21, 16, 16)
CS_FAST64_INSTANTIATE(21, 32, 16)
CS_FAST64_INSTANTIATE(37, 16, 16)
CS_FAST64_INSTANTIATE(37, 32, 16)
CS_FAST64_INSTANTIATE(53, 16, 16)
CS_FAST64_INSTANTIATE(53, 32, 16)
