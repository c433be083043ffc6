/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement of the reference's f32/f64 CPU search path, used by the
 * tests, the smoke check and the CPU-baseline leg of the benchmark as the
 * checker.  The product (coherentsearch_c_amd/) never links, loads or calls
 * anything in this directory.
 *
 * What it restates (reference file:line):
 *   cs_sinpi/cs_cospi/cs_eps            src/interp.c:17-43
 *   trial_grid_rational                 src/interp.c:80-118
 *   plan_build_one / build_group_table  src/interp.c:147-281
 *   chunk_state / harmonic_fits         src/interp.c:349-405
 *   fill_row_group / fill_chunk_profiles src/interp.c:608-861
 *   harmonic_change_trials/chunk_starts src/interp.c:866-966
 *   fold_decims                         src/search.c:294-400
 *   boxcar_widths / ladder              src/boxcar.c:78-133
 *   block_sigma / median                src/boxcar.c:18-73, 138-167
 *   analytic_sigma / boxcar_shape       src/boxcar.c:169-198
 *   psum/scan (scalar boxcar metric)    src/boxcar.c:203-252
 *   refresh_shape (count-keyed cache)   src/search.c:36-46
 *   decim_pass / chunk_after_fill       src/search.c:605-747, 830-945
 *   search_region round-robin chunks    src/search.c:989-1095
 *   search_file grid (r_lo/lobin quirk) src/search.c:1257-1266
 *   remove_duplicates/remove_harmonics  src/candidate.c:98-161
 *   candidate_profile/measure_ducy      src/candidate.c:208-329
 *   boxcar_best_width                   src/boxcar.c:736-785
 *   fourier_interp_complex              src/interp.c:317-344
 *   write_candidates / cs_format_double src/candidate.c:336-379
 *
 * The c2r transform is the project's own mixed-radix algorithm (FFTW is a
 * third-party dependency that is not in the image); its exact definition is
 * in oracle.c above orc_c2r_f32 and the HIP kernels implement the same
 * operation sequence.
 *
 * Parity pins for every function are listed in the oracle section of the
 * project's design notes and exercised by tests/test_oracle*.py.
 */
#ifndef CS_ORACLE_H
#define CS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double freq;
    double metric;
    double r;
    int nharm;
    double ducy;
} OrcCand;

typedef struct {
    int nharms;
    int m;
    double hidr;
    int ndecim;
    int decimations[64];
    int precision;          /* 0 f32, 1 f64 */
    int sigma;              /* 0 analytic, 1 measured */
    double fsp;
    double maxfrac;
    double lofreq;
    double hifreq;
    int lobin;
    int blocksize;
} OrcConfig;

typedef struct OrcCtx OrcCtx;

void orc_config_defaults(OrcConfig *c);
int  orc_decimation_set(int nharms, int maxdecim, int *out, int outcap);

/* Context: plans + chunk table for one spectrum.  amps is borrowed. */
OrcCtx *orc_ctx_create(const float *amps, long long namps, long long N, double T,
                       const OrcConfig *cfg);
void orc_ctx_free(OrcCtx *c);
long long orc_ctx_total(const OrcCtx *c);
long long orc_ctx_nchunks(const OrcCtx *c);
int  orc_ctx_chunk(const OrcCtx *c, long long ci, long long *i0, int *n, double *rstart);
double orc_ctx_r_lo(const OrcCtx *c);
double orc_ctx_lodr(const OrcCtx *c);
int  orc_ctx_rational(const OrcCtx *c, int *p, int *q);
/* number of decimation slots > 1 and their k / Hk */
int  orc_ctx_ndec(const OrcCtx *c, int *ks, int *hks);
/* sum over h of |[first fits, first unfits)| clipped to [0,total) */
long long orc_ctx_harmonic_trials(const OrcCtx *c);
/* plan scalars for harmonic h (1-based): rfloor0, dr0, P, s, base_adv */
int  orc_plan_info(const OrcCtx *c, int h, int *rfloor0, double *dr0, int *P,
                   int *s, int *base_adv);
/* W / A tables for harmonic h (W: P*m floats, A: P complex) */
int  orc_plan_tables(const OrcCtx *c, int h, float *W, float *A, uint8_t *carry,
                     int32_t *row);
/* fit predicate for harmonic h at trial t */
int  orc_harmonic_fits(const OrcCtx *c, int h, long long t);

/* Interpolated spectrum of chunk ci, h-major complex: out[(h*n + j)*2 + {0,1}],
 * h = 0..nharms (DC row and unfilled rows zero).  filled gets nharms flags.
 * form 0 = V=32 group tables (reference fill_row_group); 1 = direct per trial. */
int  orc_chunk_stack(const OrcCtx *c, long long ci, int form, float *out,
                     uint8_t *filled);

/* Profiles of chunk ci for slot (0 base, s>0 decimation s-1): out[j*nb + i]
 * (f32 or f64 by precision).  Returns nb or -1. */
int  orc_chunk_profiles(const OrcCtx *c, long long ci, int slot, void *out);

/* Per-trial metrics of chunk ci: out[slot*n + j] (double), slot 0 base,
 * NaN for j >= nvalid of a decimation.  The count-keyed shape cache is
 * replayed as a single thread walking chunks 0..ci. */
int  orc_chunk_metrics(const OrcCtx *c, long long ci, double *out);

/* FP64 oracle metrics of chunk ci (direct sinc interpolation, direct real
 * DFT, double boxcar with exact analytic sigma and shapes). */
int  orc_chunk_metrics_fp64(const OrcCtx *c, long long ci, double *out);
/* the same for rows [j0, j0+nr) of the chunk only (NaN elsewhere); measured
 * sigma is still taken over the whole chunk's sample */
int  orc_chunk_metrics_fp64_rows(const OrcCtx *c, long long ci, int j0, int nr, double *out);

/* Full raw search (candidates above threshold, reference CPU order for
 * nthreads round-robin chunks, before collapsing).  chunk_lo/chunk_hi
 * restrict to a chunk range (hi<0: all).  *out is malloc'ed. */
long long orc_search(const OrcCtx *c, double threshold, int nthreads,
                     long long chunk_lo, long long chunk_hi, OrcCand **out);

/* Project c2r (unnormalised, imag of DC/Nyquist ignored).  X: H+1 complex
 * interleaved; x: 2H reals. */
void orc_c2r_f32(int H, const float *X, float *x);
void orc_c2r_f64(int H, const double *X, double *x);

/* Widths / noise scales */
int    orc_boxcar_widths(int nbins, double fsp, double maxfrac, int *out, int cap);
int    orc_ladder_widths(int nbins, int k, const OrcConfig *cfg, int *out, int cap);
double orc_analytic_sigma(const uint8_t *filled, int k, int Hk);
void   orc_boxcar_shape(double *shape, int nbins, const int *widths, int nw,
                        const uint8_t *filled, int k, int Hk);
float  orc_block_sigma_f32(const float *M, int nbins, int n, int cap);
float  orc_profile_boxcar_f32(const float *prof, int nbins, const int *widths,
                              int nw, float invsigma, const double *shape);
double orc_profile_boxcar_f64(const double *prof, int nbins, const int *widths,
                              int nw, double invsigma, const double *shape);
int    orc_trial_grid_rational(double lodr, int *p, int *q);
double orc_sinpi(double x);
double orc_cospi(double x);
void   orc_fourier_interp_complex(double r, const float *amps, long long namps,
                                  int m, double *re, double *im);

/* Post-processing (in place on a malloc'ed array; returns new count). */
size_t orc_remove_duplicates(OrcCand *v, size_t n, double dr_tol);
size_t orc_remove_harmonics(OrcCand *v, size_t n, int numharm, double tol);
void   orc_sort_metric_desc(OrcCand *v, size_t n);
void   orc_sort_freq(OrcCand *v, size_t n);
int    orc_harmonically_related(double r1, double r2, int numharm, double tol);
void   orc_measure_ducy(const float *amps, long long namps, long long N,
                        OrcCand *v, size_t n, double fsp, double maxfrac);
int    orc_boxcar_best_width(const double *prof, int nbins, double fsp,
                             double maxfrac, int *w, double *ducy);
/* .cohout text into buf (returns bytes needed incl. NUL). */
size_t orc_format_candidates(const OrcCand *v, size_t n, char *buf, size_t cap);
void   orc_format_double(char *buf, size_t n, double x);

/* In-memory realfft + rednoise of a float series (datpipe restatement with a
 * double-precision FFT).  out: N/2 complex.  Returns 0 on success. */
int orc_realfft_rednoise(const float *samples, long long N, double dt,
                         float *out);
/* Forward realfft only (PRESTO packing: out[0]=DC, out[1]=Nyquist). */
int orc_realfft(const float *samples, long long N, float *out);
/* the Numerical Recipes unpack alone, in place on the N/2-point FFT (N floats) */
int orc_nr_unpack(float *d, long long N);
/* PRESTO's bit-trick inverse square root with one Newton step, and its
 * quickselect median (element (n-1)/2 ascending, n <= 1024), as rednoise uses them */
float orc_presto_invsqrtf(float x);
float orc_presto_median(const float *x, int n);
/* rednoise alone (datpipe.c:143-253) of a packed realfft output in[N]. */
int orc_rednoise(const float *in, long long N, double dt, float *out);

/* Zaplist on an interleaved spectrum of namps bins (src/zap.c:77-118,
 * 155-187): birds (freq_hz[i], width_hz[i], bary[i]) over T seconds of N
 * samples; returns the number of intervals zapped. */
int orc_zap(float *amps, long long namps, long long N, double T, const double *freq_hz,
            const double *width_hz, const int *bary, int nbirds, double baryv);

/* ---- PRESTO prepsubband de-dispersion (the prep front end's hot loop) ---- */
typedef struct {
    int numchan;          /* channels, lowest frequency first */
    double lofreq;        /* centre of the lowest channel, MHz */
    double chan_wid;      /* MHz */
    double dt;            /* raw sample time, s */
    int P;                /* spectra per raw block (spectra_per_subint) */
    int nsub;
    double lodm, dmstep;
    int numdms;
    int downsamp;
    long long numout;     /* per-DM output length; 0: choose_good_N(nspec / downsamp) */
    float clip_sigma;     /* time-domain clipping (clip_times) of every raw block; 0: off */
} OrcDedispConfig;

double    orc_delay_from_dm(double dm, double freq_mhz);
long long orc_choose_good_N(long long n);
/* Delay tables: dms[numdms], idispdt[numchan] (raw samples, within-subband),
 * offsets[numdms * nsub] (downsampled samples); *numout resolved.  Returns
 * blocksperread, or -1 for a configuration the reference rejects. */
int orc_dedisp_tables(const OrcDedispConfig *c, long long nspec, double *dms, int *idispdt,
                      int *offsets, long long *numout);
/* The nobary write loop over raw float spectra [nspec][numchan] (time-major):
 * out[numdms * numout] receives every DM series with its padding.  Returns the
 * number of de-dispersed samples written per DM (the rest is padding), or -1;
 * *pad receives the padding value. */
long long orc_dedisp(const OrcDedispConfig *c, const float *raw, long long nspec, float *out,
                     float *pad);
/* clip_times (clipping.c:41-215) of every raw block of raw[nspec][numchan],
 * in place, in block order, from a fresh state; levels (optional,
 * [nspec / P][numchan]) receive each block's channel levels.  Returns the
 * number of rows replaced. */
long long orc_clip_all(const OrcDedispConfig *c, float *raw, long long nspec, float *levels);
/* PRESTO avg_var (AS 52) and the lower median of n floats */
void  orc_avg_var(const float *x, int n, double *mean, double *var);
float orc_lower_median(const float *x, int n);
/* The same series by direct (unblocked) indexing: the closed form the
 * blocked loop must equal.  Writes ntime samples per DM, no padding. */
int orc_dedisp_direct(const OrcDedispConfig *c, const float *raw, long long nspec,
                      long long ntime, float *out);

#ifdef __cplusplus
}
#endif
#endif
