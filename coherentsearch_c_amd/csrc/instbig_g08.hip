/* Profile passes of the two-pass search above 64 harmonics for nharms = 73, 89, 105, 121
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(73)
CS_SPLIT_INSTANTIATE(89)
CS_SPLIT_INSTANTIATE(105)
CS_SPLIT_INSTANTIATE(121)
