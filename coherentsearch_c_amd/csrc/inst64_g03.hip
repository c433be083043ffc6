/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 4, 20, 36, 52 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(4, 16, 16)
CS_FAST64_INSTANTIATE(4, 32, 16)
CS_FAST64_INSTANTIATE(20, 16, 16)
CS_FAST64_INSTANTIATE(20, 32, 16)
CS_FAST64_INSTANTIATE(36, 16, 16)
CS_FAST64_INSTANTIATE(36, 32, 16)
CS_FAST64_INSTANTIATE(52, 16, 16)
CS_FAST64_INSTANTIATE(52, 32, 16)
