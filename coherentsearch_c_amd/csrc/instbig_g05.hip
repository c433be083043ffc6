/* Fused search kernel instantiations above 64 harmonics for nharms = 70, 86, 102, 118
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(70, 16, 16)
CS_FAST_INSTANTIATE(86, 16, 16)
CS_FAST_INSTANTIATE(102, 16, 16)
CS_FAST_INSTANTIATE(118, 16, 16)
