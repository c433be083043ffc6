/* Profile passes of the two-pass search above 64 harmonics for nharms = 69, 85, 101, 117
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(69)
CS_SPLIT_INSTANTIATE(85)
CS_SPLIT_INSTANTIATE(101)
CS_SPLIT_INSTANTIATE(117)
