/* Fused search kernel instantiations above 64 harmonics for nharms = 72, 88, 104, 120
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(72, 16, 16)
CS_FAST_INSTANTIATE(88, 16, 16)
CS_FAST_INSTANTIATE(104, 16, 16)
CS_FAST_INSTANTIATE(120, 16, 16)
