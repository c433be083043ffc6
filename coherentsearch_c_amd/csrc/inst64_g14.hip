/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 15, 31, 47, 63 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(15, 16, 16)
CS_FAST64_INSTANTIATE(15, 32, 16)
CS_FAST64_INSTANTIATE(31, 16, 16)
CS_FAST64_INSTANTIATE(31, 32, 16)
CS_FAST64_INSTANTIATE(47, 16, 16)
CS_FAST64_INSTANTIATE(47, 32, 16)
CS_FAST64_INSTANTIATE(63, 16, 16)
CS_FAST64_INSTANTIATE(63, 32, 16)
