/* Device staging (realfft + rednoise) shared between cs_stage.hip and the
 * session layer. */
#ifndef CS_STAGE_H
#define CS_STAGE_H

#include <hip/hip_runtime.h>

#include "cs_gpu.h"

/* longest rednoise block (the pipeline's endwidth) the device path holds */
#define CS_STAGE_MAXBLK 1024

extern "C" __attribute__((visibility("hidden"))) int cs_stage_enqueue(CsStage *g,
                                                                       const float *samples,
                                                                       float2 *d_out,
                                                                       hipStream_t st);

extern "C" __attribute__((visibility("hidden"))) long long cs_stage_N(const CsStage *g);

/* a stage plan on the caller's stream with stream-ordered allocations */
extern "C" __attribute__((visibility("hidden"))) CsStage *
cs_stage_create_on(long long N, double dt, int startwidth, int endwidth, double endfreq,
                   int device, hipStream_t ext);

/* a session on the caller's stream with stream-ordered allocations (ext
 * non-null), or with its own stream (ext null: cs_gpu_session_create) */
__attribute__((visibility("hidden"))) GpuSession *
cs_session_create_on(const FFTFile *ft, const SearchParams *p, const DirectPlanSet *dplans,
                     const long long *cstarts, long long total, int Nprof, long long nchunks,
                     double r_lo, double lodr, int device, hipStream_t ext);

/* sort by S/N, keep the first ncands (all when negative) and measure their
 * duty cycles on the device from the session's resident spectrum; a new list,
 * empty with a message on stderr on failure */
__attribute__((visibility("hidden"))) CandidateList
cs_session_finalize(GpuSession *s, const CandidateList *in, int ncands);

#endif
