/* Internal declarations shared by the host library's translation units. */
#ifndef CS_HOST_INTERNAL_H
#define CS_HOST_INTERNAL_H

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cands.h"
#include "core.h"
#include "cs_host.h"

/* Complex FFT of any length n in place (double, interleaved re/im),
 * sign -1 forward, +1 inverse (unnormalised).  0 on success. */
int cs_fft_any(double *z, long long n, int sign);

#endif
