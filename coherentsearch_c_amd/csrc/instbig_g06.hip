/* Profile passes of the two-pass search above 64 harmonics for nharms = 71, 87, 103, 119
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(71)
CS_SPLIT_INSTANTIATE(87)
CS_SPLIT_INSTANTIATE(103)
CS_SPLIT_INSTANTIATE(119)
