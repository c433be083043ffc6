/* Fused search kernel instantiations for nharms = 2, 18, 34, 50 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(2, 16, 16)
CS_FAST_INSTANTIATE(2, 32, 16)
CS_FAST_INSTANTIATE(18, 16, 16)
CS_FAST_INSTANTIATE(18, 32, 16)
CS_FAST_INSTANTIATE(34, 16, 16)
CS_FAST_INSTANTIATE(34, 32, 16)
CS_FAST_INSTANTIATE(50, 16, 16)
CS_FAST_INSTANTIATE(50, 32, 16)
