/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 9, 25, 41, 57 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(9, 16, 16)
CS_FAST64_INSTANTIATE(9, 32, 16)
CS_FAST64_INSTANTIATE(25, 16, 16)
CS_FAST64_INSTANTIATE(25, 32, 16)
CS_FAST64_INSTANTIATE(41, 16, 16)
CS_FAST64_INSTANTIATE(41, 32, 16)
CS_FAST64_INSTANTIATE(57, 16, 16)
CS_FAST64_INSTANTIATE(57, 32, 16)
