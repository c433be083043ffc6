/* Fused search kernel instantiations for nharms = 8, 24, 40, 56 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(8, 16, 16)
CS_FAST_INSTANTIATE(8, 32, 16)
CS_FAST_INSTANTIATE(24, 16, 16)
CS_FAST_INSTANTIATE(24, 32, 16)
CS_FAST_INSTANTIATE(40, 16, 16)
CS_FAST_INSTANTIATE(40, 32, 16)
CS_FAST_INSTANTIATE(56, 16, 16)
CS_FAST_INSTANTIATE(56, 32, 16)
