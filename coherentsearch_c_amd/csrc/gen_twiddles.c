/*
 * Build step: writes the project c2r's twiddle table as compile-time
 * constants for the device kernels (obj/cs_twiddles.h).
 *
 * The table is the one the session code computes at run time: entry
 * N(N-1)/2 + j, 1 <= N < 257, j < N, holds cos and sin of 2*pi*j/N, each a
 * separate libm call in double (the c2r's twiddle definition), as floats for
 * the f32 profiles and as doubles for the f64 ones.  Indexed with constants,
 * the kernels fold them into scalar moves instead of loading them; the
 * session code compares this table with the one it computes at run time and
 * refuses to search if the two differ.
 */
#include <math.h>
#include <stdio.h>

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif

/* through pointers the compiler cannot merge into one sincos call */
static double (*volatile cos_fn)(double) = cos;
static double (*volatile sin_fn)(double) = sin;

int main(void)
{
    printf("/* generated by gen_twiddles.c: cos, sin of 2*pi*j/N at N(N-1)/2 + j */\n");
    printf("#pragma once\n");
    printf("#define CS_TWC_NMAX 257\n");
    printf("constexpr float kTwF[] = {\n");
    for (int n = 1; n < 257; n++)
        for (int j = 0; j < n; j++) {
            const double a = 2.0 * M_PI * (double)j / (double)n;
            printf("%a,%a,\n", (double)(float)cos_fn(a), (double)(float)sin_fn(a));
        }
    printf("};\n");
    printf("constexpr double kTwD[] = {\n");
    for (int n = 1; n < 257; n++)
        for (int j = 0; j < n; j++) {
            const double a = 2.0 * M_PI * (double)j / (double)n;
            printf("%a,%a,\n", cos_fn(a), sin_fn(a));
        }
    printf("};\n");
    return 0;
}
