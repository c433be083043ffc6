/* Fused search kernel instantiations above 64 harmonics for nharms = 65, 81, 97, 113
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(65, 16, 16)
CS_FAST_INSTANTIATE(81, 16, 16)
CS_FAST_INSTANTIATE(97, 16, 16)
CS_FAST_INSTANTIATE(113, 16, 16)
