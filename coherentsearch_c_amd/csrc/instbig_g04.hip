/* Fused search kernel instantiations above 64 harmonics for nharms = 69, 85, 101, 117
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(69, 16, 16)
CS_FAST_INSTANTIATE(85, 16, 16)
CS_FAST_INSTANTIATE(101, 16, 16)
CS_FAST_INSTANTIATE(117, 16, 16)
