/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 6, 22, 38, 54 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(6, 16, 16)
CS_FAST64_INSTANTIATE(6, 32, 16)
CS_FAST64_INSTANTIATE(22, 16, 16)
CS_FAST64_INSTANTIATE(22, 32, 16)
CS_FAST64_INSTANTIATE(38, 16, 16)
CS_FAST64_INSTANTIATE(38, 32, 16)
CS_FAST64_INSTANTIATE(54, 16, 16)
CS_FAST64_INSTANTIATE(54, 32, 16)
