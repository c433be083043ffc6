/* f64-profile search kernel instantiations (--precision f64) for nharms =
 * 14, 30, 46, 62 (M = 16, 32 padded taps; decimations k <= 16). */
#include "fast_kernel.hpp"

CS_FAST64_INSTANTIATE(14, 16, 16)
CS_FAST64_INSTANTIATE(14, 32, 16)
CS_FAST64_INSTANTIATE(30, 16, 16)
CS_FAST64_INSTANTIATE(30, 32, 16)
CS_FAST64_INSTANTIATE(46, 16, 16)
CS_FAST64_INSTANTIATE(46, 32, 16)
CS_FAST64_INSTANTIATE(62, 16, 16)
CS_FAST64_INSTANTIATE(62, 32, 16)
