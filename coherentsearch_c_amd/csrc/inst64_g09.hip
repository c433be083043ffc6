/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 10, 26, 42, 58 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(10, 16, 16)
CS_FAST64_INSTANTIATE(10, 32, 16)
CS_FAST64_INSTANTIATE(26, 16, 16)
CS_FAST64_INSTANTIATE(26, 32, 16)
CS_FAST64_INSTANTIATE(42, 16, 16)
CS_FAST64_INSTANTIATE(42, 32, 16)
CS_FAST64_INSTANTIATE(58, 16, 16)
CS_FAST64_INSTANTIATE(58, 32, 16)
