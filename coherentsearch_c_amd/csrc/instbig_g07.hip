/* Profile passes of the two-pass search above 64 harmonics for nharms = 72, 88, 104, 120
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(72)
CS_SPLIT_INSTANTIATE(88)
CS_SPLIT_INSTANTIATE(104)
CS_SPLIT_INSTANTIATE(120)
