/* Profile passes of the two-pass search above 64 harmonics for nharms = 70, 86, 102, 118
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(70)
CS_SPLIT_INSTANTIATE(86)
CS_SPLIT_INSTANTIATE(102)
CS_SPLIT_INSTANTIATE(118)
