/* Profile passes of the two-pass search above 64 harmonics for nharms = 65, 81, 97, 113
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(65)
CS_SPLIT_INSTANTIATE(81)
CS_SPLIT_INSTANTIATE(97)
CS_SPLIT_INSTANTIATE(113)
