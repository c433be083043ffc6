/* Profile passes of the two-pass search above 64 harmonics for nharms = 67, 83, 99, 115
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(67)
CS_SPLIT_INSTANTIATE(83)
CS_SPLIT_INSTANTIATE(99)
CS_SPLIT_INSTANTIATE(115)
