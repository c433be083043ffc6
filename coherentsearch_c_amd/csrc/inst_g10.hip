/* Fused search kernel instantiations for nharms = 11, 27, 43, 59 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(11, 16, 16)
CS_FAST_INSTANTIATE(11, 32, 16)
CS_FAST_INSTANTIATE(27, 16, 16)
CS_FAST_INSTANTIATE(27, 32, 16)
CS_FAST_INSTANTIATE(43, 16, 16)
CS_FAST_INSTANTIATE(43, 32, 16)
CS_FAST_INSTANTIATE(59, 16, 16)
CS_FAST_INSTANTIATE(59, 32, 16)
