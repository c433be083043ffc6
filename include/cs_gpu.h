/*
 * libcs_hip.so -- the MI355X (gfx950) backend of the coherent harmonic search.
 *
 * Drop-in for the reference's CUDA backend object (src/cuda.cu): the first
 * six entry points have the reference's exact signatures and semantics, so a
 * host built against include/coherent_search.h links this library in place
 * of cuda.o / gpu_stub.o.  All structs are the reference's layouts
 * (cs_types.h).  No torch types, no C++ in the signatures.
 *
 * Differences a caller can observe:
 *   - cs_gpu_session_create returns NULL (with a message on stderr) for a
 *     configuration this build does not implement, instead of exit(1);
 *   - the analytic/measured noise scale and f32/f64 precision follow the
 *     reference's CPU path (the reference GPU path forbids measured sigma).
 */
#ifndef CS_GPU_H
#define CS_GPU_H

#include "cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define CS_API __attribute__((visibility("default")))
#else
#define CS_API
#endif

typedef struct GpuSession GpuSession;

/* Replaces src/cuda.cu:488-528 (declared include/coherent_search.h:438).
 * 1: device present (errbuf gets "name (device d, gfx950)"), -1: no device /
 * bad index / runtime error.  This build always has GPU support, so it never
 * returns 0 (the reference's gpu_stub.c:7-14 case). */
CS_API int cs_gpu_available(int device, char *errbuf, size_t errlen);

/* Replaces src/cuda.cu:554-792 (coherent_search.h:442-446).  Builds the
 * device-resident plan arena, chunk and tile tables for one observation
 * geometry.  dplans must be rational (every P > 0); p is borrowed for the
 * session's lifetime and compared by pointer in _matches, like the
 * reference.  Returns NULL on failure. */
CS_API GpuSession *cs_gpu_session_create(const FFTFile *ft, const SearchParams *p,
                                         const DirectPlanSet *dplans,
                                         const long long *cstarts, long long total,
                                         int Nprof, long long nchunks,
                                         double r_lo, double lodr, int device);

/* Replaces src/cuda.cu:1257-1266 (coherent_search.h:447-449). */
CS_API int cs_gpu_session_matches(const GpuSession *s, const FFTFile *ft,
                                  const SearchParams *p, int Nprof, long long total,
                                  double r_lo, double lodr, int device);

/* Replaces src/cuda.cu:793-1176 (coherent_search.h:450-451).  Uploads
 * ft->amps and returns every trial with metric > threshold: chunk order, then
 * base before decimations in SearchParams order, ascending trial; ducy NaN.
 * The list's v is malloc'ed; free it with free() / candlist_free().  An
 * N/namps mismatch returns an empty list, as in the reference. */
CS_API CandidateList cs_gpu_session_run(GpuSession *s, const FFTFile *ft,
                                        double threshold);

/* Replaces src/cuda.cu:1178-1235 (coherent_search.h:452). */
CS_API void cs_gpu_session_free(GpuSession *s);

/* Replaces src/cuda.cu:1238-1255 (coherent_search.h:454-460); r_hi unused. */
CS_API CandidateList cs_gpu_search_region(const FFTFile *ft, const SearchParams *p,
                                          const DirectPlanSet *dplans,
                                          const long long *cstarts, long long total,
                                          int Nprof, long long nchunks,
                                          double r_lo, double r_hi, double lodr,
                                          double threshold, int device);

/* ---- extensions (no counterpart in the reference) ---------------------- */

/* Interpolated harmonic stack of trials [t0, t0+n) after the last upload of
 * ft->amps: out[(t - t0)*(nharms+1) + h] as interleaved float pairs, h = 0
 * (DC, zero) .. nharms, zero for harmonics that do not fit the trial's chunk.
 * For parity tests against the CPU restatement. 0 on success. */
CS_API int cs_gpu_session_dump_stack(GpuSession *s, const FFTFile *ft,
                                     long long t0, long long n, float *out);

/* Upload ft->amps into the session's resident spectrum (the first half of
 * cs_gpu_session_run).  0 on success, -1 on an N/namps mismatch or error. */
CS_API int cs_gpu_session_upload(GpuSession *s, const FFTFile *ft);

/* Search the resident spectrum (the second half of cs_gpu_session_run):
 * the same candidate list without the host-to-device copy. */
CS_API CandidateList cs_gpu_session_search_resident(GpuSession *s, double threshold);

/* Device time of the last cs_gpu_session_run's search kernel(s), in ms,
 * measured with HIP events on the session's stream. */
CS_API double cs_gpu_session_last_kernel_ms(const GpuSession *s);

/* Work of one session: filled harmonic-trials (sum over harmonics of the
 * trials where the harmonic fits), trials, and the algorithmic HBM bytes of
 * the reference GPU dataflow, per trial
 *   16(nh+1) + 8 nh (2 + [any divisor decimation])
 *   + sum over non-divisor decimations k of (8(Hk+1) + 16 Hk)
 * plus 8*namps for the spectrum. */
CS_API int cs_gpu_session_work(const GpuSession *s, long long *harmonic_trials,
                               long long *trials, double *alg_bytes);

/* ------------------------------------------------------------------ */
/* Device staging of raw series (no reference counterpart on the GPU: the
 * reference stages on the host, src/datpipe.c:52-253 -- realfft_forward and
 * rednoise_apply).  A stage plan holds the FFT plan and the rednoise block
 * layout of one series length N at sample time dt, with datpipe_create's
 * startwidth/endwidth/endfreq (0 selects 6, 100, 6 Hz). */
typedef struct CsStage CsStage;

/* NULL (with a message) if N is odd, N/2 has a prime factor above 7, or the
 * rednoise widths exceed the device limit. */
CS_API CsStage *cs_gpu_stage_create(long long N, double dt, int startwidth, int endwidth,
                                    double endfreq, int device);
CS_API void cs_gpu_stage_free(CsStage *g);

/* samples (host, N floats) -> rednoised spectrum amps (host, N floats = N/2
 * complex, bin 0 = (1, 0)); packed (optional, N floats) receives the realfft
 * output before rednoise.  0 on success. */
CS_API int cs_gpu_stage_run_host(CsStage *g, const float *samples, float *amps, float *packed);

/* Stage samples (host, N floats) straight into the session's resident
 * spectrum (the device-side equivalent of datpipe_redfft_mem followed by
 * cs_gpu_session_upload); amps_host (optional, N floats) receives a copy for
 * host post-processing.  -1 if the plan's N does not match the session. */
CS_API int cs_gpu_session_stage(GpuSession *s, CsStage *g, const float *samples,
                                float *amps_host);

#ifdef __cplusplus
}
#endif
#endif
