/* Fused search kernel instantiations for nharms = 10, 26, 42, 58 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(10, 16, 16)
CS_FAST_INSTANTIATE(10, 32, 16)
CS_FAST_INSTANTIATE(26, 16, 16)
CS_FAST_INSTANTIATE(26, 32, 16)
CS_FAST_INSTANTIATE(42, 16, 16)
CS_FAST_INSTANTIATE(42, 32, 16)
CS_FAST_INSTANTIATE(58, 16, 16)
CS_FAST_INSTANTIATE(58, 32, 16)
