/* Fused search kernel instantiations above 64 harmonics for nharms = 67, 83, 99, 115
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(67, 16, 16)
CS_FAST_INSTANTIATE(83, 16, 16)
CS_FAST_INSTANTIATE(99, 16, 16)
CS_FAST_INSTANTIATE(115, 16, 16)
