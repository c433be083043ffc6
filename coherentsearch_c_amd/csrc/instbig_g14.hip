/* Profile passes of the two-pass search above 64 harmonics for nharms = 79, 95, 111, 127
 * (f32 profiles, decimations k <= 16).  Harmonic counts are interleaved
 * across the group files so that their compile times balance. */
#include "split_kernel.hpp"

CS_SPLIT_INSTANTIATE(79)
CS_SPLIT_INSTANTIATE(95)
CS_SPLIT_INSTANTIATE(111)
CS_SPLIT_INSTANTIATE(127)
