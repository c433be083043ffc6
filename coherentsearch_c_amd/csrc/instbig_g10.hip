/* Fused search kernel instantiations above 64 harmonics for nharms = 75, 91, 107, 123
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(75, 16, 16)
CS_FAST_INSTANTIATE(91, 16, 16)
CS_FAST_INSTANTIATE(107, 16, 16)
CS_FAST_INSTANTIATE(123, 16, 16)
