/* Fused search kernel instantiations above 64 harmonics for nharms = 66, 82, 98, 114
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(66, 16, 16)
CS_FAST_INSTANTIATE(82, 16, 16)
CS_FAST_INSTANTIATE(98, 16, 16)
CS_FAST_INSTANTIATE(114, 16, 16)
