/* Profile passes of the two-pass search above 64 harmonics for nharms = 75, 91, 107, 123
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(75)
CS_SPLIT_INSTANTIATE(91)
CS_SPLIT_INSTANTIATE(107)
CS_SPLIT_INSTANTIATE(123)
