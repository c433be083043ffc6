/* Fused search kernel instantiations for nharms = 16, 32, 48, 64 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(16, 16, 16)
CS_FAST_INSTANTIATE(16, 32, 16)
CS_FAST_INSTANTIATE(32, 16, 16)
CS_FAST_INSTANTIATE(32, 32, 16)
CS_FAST_INSTANTIATE(48, 16, 16)
CS_FAST_INSTANTIATE(48, 32, 16)
CS_FAST_INSTANTIATE(64, 16, 16)
CS_FAST_INSTANTIATE(64, 32, 16)
