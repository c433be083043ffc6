/* Fused search kernel instantiations above 64 harmonics for nharms = 74, 90, 106, 122
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(74, 16, 16)
CS_FAST_INSTANTIATE(90, 16, 16)
CS_FAST_INSTANTIATE(106, 16, 16)
CS_FAST_INSTANTIATE(122, 16, 16)
