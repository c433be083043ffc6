/* Fused search kernel instantiations above 64 harmonics for nharms = 80, 96, 112, 128
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(80, 16, 16)
CS_FAST_INSTANTIATE(96, 16, 16)
CS_FAST_INSTANTIATE(112, 16, 16)
CS_FAST_INSTANTIATE(128, 16, 16)
