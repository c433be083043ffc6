/* Profile passes of the two-pass search above 64 harmonics for nharms = 80, 96, 112, 128
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(80)
CS_SPLIT_INSTANTIATE(96)
CS_SPLIT_INSTANTIATE(112)
CS_SPLIT_INSTANTIATE(128)
