/* Fused search kernel instantiations above 64 harmonics for nharms = 77, 93, 109, 125
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(77, 16, 16)
CS_FAST_INSTANTIATE(93, 16, 16)
CS_FAST_INSTANTIATE(109, 16, 16)
CS_FAST_INSTANTIATE(125, 16, 16)
