/* Fused search kernel instantiations for nharms = 12, 28, 44, 60 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(12, 16, 16)
CS_FAST_INSTANTIATE(12, 32, 16)
CS_FAST_INSTANTIATE(28, 16, 16)
CS_FAST_INSTANTIATE(28, 32, 16)
CS_FAST_INSTANTIATE(44, 16, 16)
CS_FAST_INSTANTIATE(44, 32, 16)
CS_FAST_INSTANTIATE(60, 16, 16)
CS_FAST_INSTANTIATE(60, 32, 16)
