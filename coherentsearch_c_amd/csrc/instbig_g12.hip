/* Profile passes of the two-pass search above 64 harmonics for nharms = 77, 93, 109, 125
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(77)
CS_SPLIT_INSTANTIATE(93)
CS_SPLIT_INSTANTIATE(109)
CS_SPLIT_INSTANTIATE(125)
