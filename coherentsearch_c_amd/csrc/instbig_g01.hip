/* Profile passes of the two-pass search above 64 harmonics for nharms = 66, 82, 98, 114
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(66)
CS_SPLIT_INSTANTIATE(82)
CS_SPLIT_INSTANTIATE(98)
CS_SPLIT_INSTANTIATE(114)
