/* Profile passes of the two-pass search above 64 harmonics for nharms = 78, 94, 110, 126
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(78)
CS_SPLIT_INSTANTIATE(94)
CS_SPLIT_INSTANTIATE(110)
CS_SPLIT_INSTANTIATE(126)
