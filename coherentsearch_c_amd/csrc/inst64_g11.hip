/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 12, 28, 44, 60 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(12, 16, 16)
CS_FAST64_INSTANTIATE(12, 32, 16)
CS_FAST64_INSTANTIATE(28, 16, 16)
CS_FAST64_INSTANTIATE(28, 32, 16)
CS_FAST64_INSTANTIATE(44, 16, 16)
CS_FAST64_INSTANTIATE(44, 32, 16)
CS_FAST64_INSTANTIATE(60, 16, 16)
CS_FAST64_INSTANTIATE(60, 32, 16)
