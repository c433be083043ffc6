/* Fused search kernel instantiations for nharms = 5, 21, 37, 53 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(5, 16, 16)
CS_FAST_INSTANTIATE(5, 32, 16)
CS_FAST_INSTANTIATE(21, 16, 16)
CS_FAST_INSTANTIATE(21, 32, 16)
CS_FAST_INSTANTIATE(37, 16, 16)
CS_FAST_INSTANTIATE(37, 32, 16)
CS_FAST_INSTANTIATE(53, 16, 16)
CS_FAST_INSTANTIATE(53, 32, 16)
