/* Fused search kernel instantiations for nharms = 13, 29, 45, 61 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(13, 16, 16)
CS_FAST_INSTANTIATE(13, 32, 16)
CS_FAST_INSTANTIATE(29, 16, 16)
CS_FAST_INSTANTIATE(29, 32, 16)
CS_FAST_INSTANTIATE(45, 16, 16)
CS_FAST_INSTANTIATE(45, 32, 16)
CS_FAST_INSTANTIATE(61, 16, 16)
CS_FAST_INSTANTIATE(61, 32, 16)
