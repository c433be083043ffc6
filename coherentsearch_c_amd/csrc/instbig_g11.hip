/* Profile passes of the two-pass search above 64 harmonics for nharms = 76, 92, 108, 124
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(76)
CS_SPLIT_INSTANTIATE(92)
CS_SPLIT_INSTANTIATE(108)
CS_SPLIT_INSTANTIATE(124)
