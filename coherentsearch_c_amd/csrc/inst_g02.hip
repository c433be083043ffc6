/* Fused search kernel instantiations for nharms = 3, 19, 35, 51 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(3, 16, 16)
CS_FAST_INSTANTIATE(3, 32, 16)
CS_FAST_INSTANTIATE(19, 16, 16)
CS_FAST_INSTANTIATE(19, 32, 16)
CS_FAST_INSTANTIATE(35, 16, 16)
CS_FAST_INSTANTIATE(35, 32, 16)
CS_FAST_INSTANTIATE(51, 16, 16)
CS_FAST_INSTANTIATE(51, 32, 16)
