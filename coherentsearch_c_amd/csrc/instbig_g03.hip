/* Profile passes of the two-pass search above 64 harmonics for nharms = 68, 84, 100, 116
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(68)
CS_SPLIT_INSTANTIATE(84)
CS_SPLIT_INSTANTIATE(100)
CS_SPLIT_INSTANTIATE(116)
