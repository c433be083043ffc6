/* Fused search kernel instantiations above 64 harmonics for nharms = 73, 89, 105, 121
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(73, 16, 16)
CS_FAST_INSTANTIATE(89, 16, 16)
CS_FAST_INSTANTIATE(105, 16, 16)
CS_FAST_INSTANTIATE(121, 16, 16)
