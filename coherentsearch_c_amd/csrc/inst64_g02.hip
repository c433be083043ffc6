/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 3, 19, 35, 51 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(3, 16, 16)
CS_FAST64_INSTANTIATE(3, 32, 16)
CS_FAST64_INSTANTIATE(19, 16, 16)
CS_FAST64_INSTANTIATE(19, 32, 16)
CS_FAST64_INSTANTIATE(35, 16, 16)
CS_FAST64_INSTANTIATE(35, 32, 16)
CS_FAST64_INSTANTIATE(51, 16, 16)
CS_FAST64_INSTANTIATE(51, 32, 16)
