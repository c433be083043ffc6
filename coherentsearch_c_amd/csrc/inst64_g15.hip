/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 16, 32, 48, 64 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(16, 16, 16)
CS_FAST64_INSTANTIATE(16, 32, 16)
CS_FAST64_INSTANTIATE(32, 16, 16)
CS_FAST64_INSTANTIATE(32, 32, 16)
CS_FAST64_INSTANTIATE(48, 16, 16)
CS_FAST64_INSTANTIATE(48, 32, 16)
CS_FAST64_INSTANTIATE(64, 16, 16)
CS_FAST64_INSTANTIATE(64, 32, 16)
