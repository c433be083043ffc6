/* Fused search kernel instantiations above 64 harmonics for nharms = 76, 92, 108, 124
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(76, 16, 16)
CS_FAST_INSTANTIATE(92, 16, 16)
CS_FAST_INSTANTIATE(108, 16, 16)
CS_FAST_INSTANTIATE(124, 16, 16)
