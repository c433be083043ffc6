/* Fused search kernel instantiations above 64 harmonics for nharms = 79, 95, 111, 127
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(79, 16, 16)
CS_FAST_INSTANTIATE(95, 16, 16)
CS_FAST_INSTANTIATE(111, 16, 16)
CS_FAST_INSTANTIATE(127, 16, 16)
