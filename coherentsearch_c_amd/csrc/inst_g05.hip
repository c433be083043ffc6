/* Fused search kernel instantiations for nharms = 6, 22, 38, 54 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(6, 16, 16)
CS_FAST_INSTANTIATE(6, 32, 16)
CS_FAST_INSTANTIATE(22, 16, 16)
CS_FAST_INSTANTIATE(22, 32, 16)
CS_FAST_INSTANTIATE(38, 16, 16)
CS_FAST_INSTANTIATE(38, 32, 16)
CS_FAST_INSTANTIATE(54, 16, 16)
CS_FAST_INSTANTIATE(54, 32, 16)
