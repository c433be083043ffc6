/* Profile passes of the two-pass search above 64 harmonics for nharms = 74, 90, 106, 122
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(74)
CS_SPLIT_INSTANTIATE(90)
CS_SPLIT_INSTANTIATE(106)
CS_SPLIT_INSTANTIATE(122)
