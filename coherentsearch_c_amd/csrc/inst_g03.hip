/* Fused search kernel instantiations for nharms = 4, 20, 36, 52 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(4, 16, 16)
CS_FAST_INSTANTIATE(4, 32, 16)
CS_FAST_INSTANTIATE(20, 16, 16)
CS_FAST_INSTANTIATE(20, 32, 16)
CS_FAST_INSTANTIATE(36, 16, 16)
CS_FAST_INSTANTIATE(36, 32, 16)
CS_FAST_INSTANTIATE(52, 16, 16)
CS_FAST_INSTANTIATE(52, 32, 16)
