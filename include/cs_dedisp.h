/*
 * De-dispersion of raw filterbank spectra into DM series resident in HBM:
 * the hot loop of the reference's prep front end (PRESTO prepsubband run in
 * memory, src/prep_dedisp.c:83-892 with get_data at :1104-1427), on the GPU.
 *
 * The reference de-disperses block by block (get_data, read_subbands_GPU,
 * downsamp_GPU, float_dedisp_GPU), copies every DM's series back to the host
 * and hands them to prep_search_run (src/prep_search.c:44-145), which stages
 * and searches each one.  Here the series stay on the device: a search
 * request of the engine (cs_gpu.h, CsSearchRequest.samples) takes a pointer
 * into the plan's series directly.
 *
 * Scope: the nobary path with or without clipping (-clip / -noclip), no
 * masks, zero-DM or ignored channels (prepsubband -nobary), on spectra already read into
 * memory: nspec x numchan, time-major, lowest frequency first, float32 (what
 * PRESTO's get_rawblock returns) or uint8 (8-bit filterbank samples as their
 * integer values).  The series are bit-identical to the reference's CPU and
 * GPU paths: the same delay tables (double), the same float sums in the same
 * order, the same written length and the same padding.
 */
#ifndef CS_DEDISP_H
#define CS_DEDISP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef CS_API
#define CS_API __attribute__((visibility("default")))
#endif

/* prepsubband's options and the observation's header values
 * (src/prep_dedisp.c:320-331, 362-367, 405-422) */
typedef struct {
    int numchan;            /* channels (s->num_channels) */
    double lofreq;          /* centre frequency of the lowest channel, MHz (idata.freq) */
    double chan_wid;        /* channel width, MHz (> 0: ascending order) */
    double dt;              /* raw sample time, s */
    int spectra_per_subint; /* raw block length the reader returns */
    int nsub;               /* -nsub (numchan must be a multiple) */
    double lodm, dmstep;    /* -lodm, -dmstep */
    int numdms;             /* -numdms */
    int downsamp;           /* -downsamp (a factor of spectra_per_subint) */
    long long numout;       /* -numout; 0: choose_good_N(nspec / downsamp) */
    float clip_sigma;       /* -clip: time-domain clipping of every raw block
                               (clip_times, clipping.c:41-215); 0: -noclip */
} CsDedispConfig;

typedef struct {
    int blocksperread;      /* raw blocks per work block (src/prep_dedisp.c:366) */
    int worklen;            /* raw spectra per work block */
    int nsub;
    int max_offset;         /* largest subband offset, downsampled samples */
    long long numout;       /* samples per DM series (written + padding) */
    long long written;      /* de-dispersed samples per DM; known at create */
    long long chunk;        /* raw spectra per device pass */
    double dsdt;            /* output sample time, s */
    float pad;              /* padding value (after a run: the running mean of DM 0) */
    int pad_;
} CsDedispInfo;

typedef struct CsDedisp CsDedisp;

/* A plan for nspec spectra (a multiple of spectra_per_subint) on `device`:
 * the delay tables, the device rings and the resident series
 * [numdms][numout] floats.  NULL (message on stderr) for a configuration the
 * reference rejects: nsub not dividing numchan, downsamp not dividing the
 * block, more dispersion across a subband than one block
 * (backend_common.c:1706-1712), negative delays.  device -1: the tables
 * and info only (no device resources; cs_gpu_dedisp_run refuses). */
CS_API CsDedisp *cs_gpu_dedisp_create(const CsDedispConfig *c, long long nspec, int device);
CS_API void cs_gpu_dedisp_free(CsDedisp *d);
CS_API int cs_gpu_dedisp_info(const CsDedisp *d, CsDedispInfo *out);
/* dms[numdms], idispdt[numchan] (raw samples), offsets[numdms * nsub]
 * (downsampled samples); any may be NULL */
CS_API int cs_gpu_dedisp_tables(const CsDedisp *d, double *dms, int *idispdt, int *offsets);

/* De-disperse `spectra` (host memory or this device's memory; float32 when
 * is_uint8 is 0) into the resident series and pad them.  Returns 0; the
 * series' device address comes from cs_gpu_dedisp_series. */
CS_API int cs_gpu_dedisp_run(CsDedisp *d, const void *spectra, int is_uint8);
/* device address of DM i's series (numout floats), NULL out of range */
CS_API const float *cs_gpu_dedisp_series(const CsDedisp *d, int i);
/* copy DMs [i0, i0 + n) to host memory (n * numout floats) */
CS_API int cs_gpu_dedisp_copy_out(const CsDedisp *d, int i0, int n, float *host);
/* milliseconds of the last run's device work (HIP events on the plan's
 * stream, from the run's start, clipping included, to its last DM pass), and
 * the subband and DM kernels' shares */
CS_API double cs_gpu_dedisp_last_ms(const CsDedisp *d, double *subband_ms, double *dm_ms);

/* PRESTO helpers the tables use (third_party/presto/src/misc_utils.c:334-373,
 * dispersion.c:100-108) */
CS_API long long cs_choose_good_N(long long n);
CS_API double cs_delay_from_dm(double dm, double freq_mhz);

#ifdef __cplusplus
}
#endif
#endif
