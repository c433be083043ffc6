/* Fused search kernel instantiations for nharms = 14, 30, 46, 62 (M = 16, 32 padded taps;
 * decimations k <= 16).  Harmonic counts are interleaved across the group
 * files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(14, 16, 16)
CS_FAST_INSTANTIATE(14, 32, 16)
CS_FAST_INSTANTIATE(30, 16, 16)
CS_FAST_INSTANTIATE(30, 32, 16)
CS_FAST_INSTANTIATE(46, 16, 16)
CS_FAST_INSTANTIATE(46, 32, 16)
CS_FAST_INSTANTIATE(62, 16, 16)
CS_FAST_INSTANTIATE(62, 32, 16)
