/* Fused search kernel instantiations for nharms = 15, 31, 47, 63 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(15, 16, 16)
CS_FAST_INSTANTIATE(15, 32, 16)
CS_FAST_INSTANTIATE(31, 16, 16)
CS_FAST_INSTANTIATE(31, 32, 16)
CS_FAST_INSTANTIATE(47, 16, 16)
CS_FAST_INSTANTIATE(47, 32, 16)
CS_FAST_INSTANTIATE(63, 16, 16)
CS_FAST_INSTANTIATE(63, 32, 16)
