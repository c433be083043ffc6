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
/* The same for long transforms, given n points of scratch (the multi-level
 * cache-blocked path from 16384 points when n is 7-smooth) on nthreads
 * OpenMP threads (0: the OpenMP default). */
int cs_fft_work(double *z, double *scratch, long long n, int sign, int nthreads);

/* Calls fn(line, len, ctx) for every line of a text file (the newline not
 * included; a last line without one counts).  0 when every call returned 0,
 * -1 if the file cannot be read or a call failed. */
int cs_for_each_line(const char *path, int (*fn)(const char *, size_t, void *), void *ctx);

#endif
