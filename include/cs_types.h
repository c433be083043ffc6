/*
 * Types that cross the coherent-search hot-path boundary.
 *
 * Every struct here keeps the byte layout of the reference's public header
 * (include/coherent_search.h) on x86-64 SysV, so a host built against the
 * reference can hand its own FFTFile / SearchParams / DirectPlanSet to this
 * library and free the CandidateList it gets back with free().  Layouts are
 * static-asserted at the bottom.  Nothing here depends on FFTW.
 *
 * Reference anchors (all in include/coherent_search.h):
 *   tunables            :34-43      SimpleInf / FFTFile  :48-76
 *   cs_precision/sigma  :115-116    SearchParams         :118-131
 *   Candidate(List)     :139-151    DirectPlan(Set)      :170-195
 *   ZapBird / ZapList   :83-93
 */
#ifndef CS_TYPES_H
#define CS_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Tunables, values as in coherent_search.h:34-43. */
#define CS_DIRECT_GROUP_V 32
#define CS_DIRECT_QMAX (1 << 16)
#define CS_BOXCAR_SIGMA_SAMPLES 8192
#define CS_BC_BATCH 128
#define CS_LADDER_WMAX 6
#define CS_CPU_DEFAULT_BLOCKSIZE 2048
#define CS_GPU_DEFAULT_BLOCKSIZE 65536
#define CS_SIGMA_CHECK_CHUNKS 3
#define CS_SIGMA_CHECK_TOL 0.10

typedef struct {
    char *path;
    char *object;      /* may be NULL */
    int   has_epoch;
    double epoch;
    long long N;       /* number of time-series bins */
    int   has_N;
    double dt;         /* time-series bin width (s) */
    int   has_dt;
    double DM;
    int   has_DM;
} SimpleInf;

typedef struct {
    char *path;
    SimpleInf inf;
    long long N;
    double T;          /* N * dt */
    double df;         /* 1 / T */
    int dereddened;
    int detrended;
    float DC;
    float Nyquist;
    float *amps;       /* 2*namps interleaved floats (re, im) */
    long long namps;
    void *map;
    size_t mapsize;
    int owns_amps;     /* free(amps) on close when set */
} FFTFile;

typedef enum { CS_PREC_F32 = 0, CS_PREC_F64 = 1 } cs_precision;
typedef enum { CS_SIGMA_ANALYTIC = 0, CS_SIGMA_MEASURED = 1 } cs_sigma;

typedef struct {
    int nharms;
    int m;
    int numbetween;
    double hidr;
    double threshold;
    double boxcar_fsp;
    double boxcar_maxfrac;
    double boxcar_gatemargin;
    int ndecim;
    int decimations[64];
    cs_precision precision;
    cs_sigma sigma;
} SearchParams;

typedef struct {
    double freq;
    double metric;
    double r;
    int nharm;
    double ducy;
} Candidate;

typedef struct {
    Candidate *v;
    size_t n;
    size_t cap;
} CandidateList;

/* One harmonic's interpolation plan.  This library reads only the scalar
 * fields and row / W / A_re / A_im / carry; the V=32 group tables (ngrp ..
 * gA_im) exist for layout compatibility with the reference's CPU filler and
 * may be NULL. */
typedef struct {
    int h;
    int m;
    int q;
    int s;
    int base_adv;
    int pnum;
    int rfloor0;
    double dr0;
    int P;
    int32_t *row;          /* q entries, 1-based phase index or 0 */
    float *W;              /* (m, P), column-major: W[p*m + i] */
    float *A_re, *A_im;    /* (P,) conj(A(dr)) */
    uint8_t *carry;        /* (P,) */
    int ngrp;
    int32_t *grow;
    int32_t *goff;
    int32_t *gnj;
    float *gW;
    float *gA_re, *gA_im;
} DirectPlan;

typedef struct {
    DirectPlan *plans;
    int n;
} DirectPlanSet;

/* One zaplist line: a frequency and width (Hz); bary = 1 for a 'B' line,
 * used as given, otherwise scaled by (1 + baryv). */
typedef struct {
    double freq;
    double width;
    int bary;
} ZapBird;

typedef struct {
    ZapBird *v;
    int n;
} ZapList;

#ifdef __cplusplus
}
#endif

/* Layout checks against the reference's x86-64 ABI. */
#ifdef __cplusplus
#define CS_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define CS_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
CS_STATIC_ASSERT(sizeof(SimpleInf) == 80, "SimpleInf layout");
CS_STATIC_ASSERT(sizeof(FFTFile) == 168, "FFTFile layout");
CS_STATIC_ASSERT(offsetof(FFTFile, N) == 88, "FFTFile.N");
CS_STATIC_ASSERT(offsetof(FFTFile, T) == 96, "FFTFile.T");
CS_STATIC_ASSERT(offsetof(FFTFile, amps) == 128, "FFTFile.amps");
CS_STATIC_ASSERT(offsetof(FFTFile, namps) == 136, "FFTFile.namps");
CS_STATIC_ASSERT(sizeof(SearchParams) == 328, "SearchParams layout");
CS_STATIC_ASSERT(offsetof(SearchParams, hidr) == 16, "SearchParams.hidr");
CS_STATIC_ASSERT(offsetof(SearchParams, ndecim) == 56, "SearchParams.ndecim");
CS_STATIC_ASSERT(offsetof(SearchParams, decimations) == 60, "SearchParams.decimations");
CS_STATIC_ASSERT(offsetof(SearchParams, precision) == 316, "SearchParams.precision");
CS_STATIC_ASSERT(offsetof(SearchParams, sigma) == 320, "SearchParams.sigma");
CS_STATIC_ASSERT(sizeof(DirectPlan) == 144, "DirectPlan layout");
CS_STATIC_ASSERT(offsetof(DirectPlan, dr0) == 32, "DirectPlan.dr0");
CS_STATIC_ASSERT(offsetof(DirectPlan, row) == 48, "DirectPlan.row");
CS_STATIC_ASSERT(offsetof(DirectPlan, carry) == 80, "DirectPlan.carry");
CS_STATIC_ASSERT(offsetof(DirectPlan, gA_im) == 136, "DirectPlan.gA_im");
CS_STATIC_ASSERT(sizeof(DirectPlanSet) == 16, "DirectPlanSet layout");
CS_STATIC_ASSERT(sizeof(Candidate) == 40, "Candidate layout");
CS_STATIC_ASSERT(offsetof(Candidate, nharm) == 24, "Candidate.nharm");
CS_STATIC_ASSERT(offsetof(Candidate, ducy) == 32, "Candidate.ducy");
CS_STATIC_ASSERT(sizeof(CandidateList) == 24, "CandidateList layout");
CS_STATIC_ASSERT(sizeof(ZapBird) == 24, "ZapBird layout");
CS_STATIC_ASSERT(sizeof(ZapList) == 16, "ZapList layout");

#endif /* CS_TYPES_H */
