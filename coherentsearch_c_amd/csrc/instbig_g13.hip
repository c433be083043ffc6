/* Fused search kernel instantiations above 64 harmonics for nharms = 78, 94, 110, 126
 * (M = 16 padded taps, f32 profiles; decimations k <= 16).  Harmonic counts are
 * interleaved across the group files so that their compile times balance. */
#include "fast_kernel.hpp"

CS_FAST_INSTANTIATE(78, 16, 16)
CS_FAST_INSTANTIATE(94, 16, 16)
CS_FAST_INSTANTIATE(110, 16, 16)
CS_FAST_INSTANTIATE(126, 16, 16)
