/* Device staging (realfft + rednoise) shared between cs_stage.hip and the
 * session layer. */
#ifndef CS_STAGE_H
#define CS_STAGE_H

#include <hip/hip_runtime.h>

#include "cs_gpu.h"

/* longest rednoise block (the pipeline's endwidth) the device path holds */
#define CS_STAGE_MAXBLK 1024

/* Table uploads through a per-thread pinned staging buffer: cs_upload_buffer
 * returns host memory for `bytes` (waiting for this thread's previous upload
 * to finish reading it), cs_upload copies its first `bytes` to dst with one
 * asynchronous DMA on st. */
__attribute__((visibility("hidden"))) void *cs_upload_buffer(size_t bytes);
__attribute__((visibility("hidden"))) hipError_t cs_upload(void *dst, size_t bytes,
                                                           hipStream_t st);

/* offsets of consecutive 256-byte aligned parts of one device block */
static inline size_t cs_part(size_t *off, size_t bytes)
{
    const size_t o = *off;
    *off = (o + bytes + 255) & ~(size_t)255;
    return o;
}

extern "C" __attribute__((visibility("hidden"))) int cs_stage_enqueue(CsStage *g,
                                                                       const float *samples,
                                                                       float2 *d_out,
                                                                       hipStream_t st);

/* the two halves of cs_stage_enqueue: the samples' copy into the plan, and the
 * staging kernels on samples already there (a caller can do host work while
 * the copy runs) */
extern "C" __attribute__((visibility("hidden"))) int cs_stage_copy_in(CsStage *g,
                                                                       const float *samples,
                                                                       hipStream_t st);
extern "C" __attribute__((visibility("hidden"))) int cs_stage_run(CsStage *g, float2 *d_out,
                                                                   hipStream_t st);

/* the same, with the staging kernels replayed from a graph captured once per
 * output buffer */
extern "C" __attribute__((visibility("hidden"))) int cs_stage_enqueue_graph(CsStage *g,
                                                                             const float *samples,
                                                                             float2 *d_out,
                                                                             hipStream_t st);

/* move a stage plan to another stream (its next operations run there) */
extern "C" __attribute__((visibility("hidden"))) void cs_stage_bind_stream(CsStage *g,
                                                                            hipStream_t st);

extern "C" __attribute__((visibility("hidden"))) long long cs_stage_N(const CsStage *g);

/* device bytes of the transform buffers of a stage plan for N samples
 * (Bluestein's three convolution buffers when N/2 has a prime factor above 7) */
extern "C" __attribute__((visibility("hidden"))) long long cs_stage_fft_bytes(long long N);

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

/* engine sessions: run on `st` from now on, replaying launch graphs when
 * `graphs` (search graph keyed on threshold and record buffer) */
__attribute__((visibility("hidden"))) void cs_session_bind_stream(GpuSession *s, hipStream_t st,
                                                                  bool graphs);
/* upload samples and stage them into the session's spectrum with the stage
 * plan's graph, on the session's stream (not synchronised) */
__attribute__((visibility("hidden"))) int cs_session_stage_graph(GpuSession *s, CsStage *g,
                                                                 const float *samples);
/* stage samples already copied into g (cs_stage_copy_in, same stream) into the
 * session's spectrum, copying it to amps_host when given; synchronised */
__attribute__((visibility("hidden"))) int cs_session_stage_loaded(GpuSession *s, CsStage *g,
                                                                  float *amps_host);

/* device bytes a session of p at trial step lodr adds for the general
 * kernel's work buffer (an upper bound; 0 when a register kernel serves it) */
__attribute__((visibility("hidden"))) long long cs_session_general_bytes(const SearchParams *p,
                                                                        double lodr);

/* launch shape of the search kernel (no effect on results) */
__attribute__((visibility("hidden"))) void cs_session_set_launch(GpuSession *s, int lane_order,
                                                                 int xcd_stripe);

/* the tuned launch of a kernel class (device, nharms, padded taps, f64,
 * measured sigma), if cs_gpu_autotune has produced or loaded one */
__attribute__((visibility("hidden"))) bool cs_tune_lookup(int device, int nh, int mp, bool f64,
                                                          bool measured, CsTuneResult *out);

#endif