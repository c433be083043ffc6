/* Fused search kernel instantiations for nharms = 9, 25, 41, 57 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(9, 16, 16)
CS_FAST_INSTANTIATE(9, 32, 16)
CS_FAST_INSTANTIATE(25, 16, 16)
CS_FAST_INSTANTIATE(25, 32, 16)
CS_FAST_INSTANTIATE(41, 16, 16)
CS_FAST_INSTANTIATE(41, 32, 16)
CS_FAST_INSTANTIATE(57, 16, 16)
CS_FAST_INSTANTIATE(57, 32, 16)
