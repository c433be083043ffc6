/* Fused search kernel instantiations above 64 harmonics for nharms = 71, 87, 103, 119
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(71, 16, 16)
CS_FAST_INSTANTIATE(87, 16, 16)
CS_FAST_INSTANTIATE(103, 16, 16)
CS_FAST_INSTANTIATE(119, 16, 16)
