/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 2, 18, 34, 50 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(2, 16, 16)
CS_FAST64_INSTANTIATE(2, 32, 16)
CS_FAST64_INSTANTIATE(18, 16, 16)
CS_FAST64_INSTANTIATE(18, 32, 16)
CS_FAST64_INSTANTIATE(34, 16, 16)
CS_FAST64_INSTANTIATE(34, 32, 16)
CS_FAST64_INSTANTIATE(50, 16, 16)
CS_FAST64_INSTANTIATE(50, 32, 16)
