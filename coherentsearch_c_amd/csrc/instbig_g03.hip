/* Fused search kernel instantiations above 64 harmonics for nharms = 68, 84, 100, 116
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(68, 16, 16)
CS_FAST_INSTANTIATE(84, 16, 16)
CS_FAST_INSTANTIATE(100, 16, 16)
CS_FAST_INSTANTIATE(116, 16, 16)
