/*
 * Option-compatible driver for coherent_search / coherent_search_list
 * (src/cli.c:351-1116): the same options and defaults, output naming
 * (<stem>.cohout per input unless a single input without --outdir, which
 * goes to -o or stdout), output-collision check, list files relative to the
 * list's directory, per-file search_file with one shared SearchCache, then
 * sort by S/N, truncate to --ncands, duty cycles and .cohout.  With more than
 * one input, producer threads prepare the next inputs while the current one
 * is searched (src/cli.c:556-686).  .dat inputs are realfft'd and
 * rednoise'd on the GPU (on the producers' own streams, beside the search)
 * unless --host-staging asks for the host transform the reference uses.  The
 * search always runs on the GPU (--gpu N picks the device, default 0),
 * including --metricstats and --normalize, which the reference runs on the
 * CPU only: their report goes to stderr and <base>_metricstats.txt,
 * _metricfap.txt and _metrichist.txt (src/cli.c:1049-1091).  The plotting
 * options are rejected with exit status 2; the rest of the CPU-only knobs are
 * accepted and ignored.
 */
#define _GNU_SOURCE
#include "host_internal.h"

#include <errno.h>
#include <omp.h>
#include <pthread.h>
#include <sys/stat.h>
#include <sys/types.h>

typedef struct {
    char **files;
    int nfiles, files_owned;
    double threshold;
    const char *outfilenm, *outdir, *zaplist;
    int nharms, ncands, lobin, m, blocksize, maxdecim, numharm;
    double lofreq, hifreq, hidr, drtol, baryv;
    cs_precision precision;
    cs_sigma sigma;
    int remove, harm_remove, verbose, gpu_device;
    int metricstats, normalize;
    int host_staging;
} Opts;

static void usage(FILE *f, int list)
{
    if (list) fprintf(f, "Usage: coherent_search_list [options] FILE.dat.list\n");
    else fprintf(f, "Usage: coherent_search [options] FILE.fft|FILE.dat [FILE2 ...]\n");
    fprintf(f,
"\nCoherent harmonic-summing pulsar search on an AMD Instinct GPU.\n"
"Options (defaults as the reference):\n"
"  -t, --threshold X   S/N cutoff (8.0)        -o, --outputfilenm F  single output\n"
"      --outdir D      per-file .cohout dir    -n, --nharms N        harmonics (60)\n"
"      --ncands N      max candidates (100)        --lobin N         (100)\n"
"      --lofreq F      (0.1 Hz)                    --hifreq F        (125 Hz)\n"
"      --hidr F        (0.5)                       --m N             taps (16)\n"
"      --precision P   f32|f64                     --sigma S         analytic|measured\n"
"      --blocksize N   trials per chunk (65536)    --maxdecim N      (6)\n"
"      --drtol F       (1.0)                       --numharm N       (16)\n"
"      --noremove      --noharmremove              --gpu [N]         device (0)\n"
"      --zaplist F     zap the listed birds (Hz, width Hz; 'B' lines barycentric)\n"
"      --baryv V       v/c applied to non-'B' zaplist lines (0.0)\n"
"      --metricstats   per-block and per-(k, frequency) metric statistics report\n"
"      --normalize     metric in noise units per (k, frequency window), two passes\n"
"      --host-staging  realfft + rednoise of .dat inputs on the host (default: on\n"
"                      the GPU)\n"
"  -v, --verbose       --threads N  --nowisdom  --wisdomfile F  --progressbar\n"
"      --noprogress    --noplot     --help\n");
}

static int parse_d(const char *s, double *v)
{
    char *e = NULL;
    errno = 0;
    const double x = strtod(s, &e);
    if (errno || !e || *e) return -1;
    *v = x;
    return 0;
}
static int parse_i(const char *s, int *v)
{
    char *e = NULL;
    errno = 0;
    const long x = strtol(s, &e, 10);
    if (errno || !e || *e) return -1;
    *v = (int)x;
    return 0;
}

/* The path without its extension: the extension is the basename's last '.'
 * and what follows, unless that '.' starts the basename (Julia's splitext,
 * the reference's output naming, src/cli.c:74-96). */
static char *path_stem(const char *path)
{
    const char *base = strrchr(path, '/');
    base = base ? base + 1 : path;
    const char *dot = strrchr(base, '.');
    const size_t keep = dot && dot > base ? (size_t)(dot - path) : strlen(path);
    return strndup(path, keep);
}

static char *output_path(const char *in, const char *outfilenm, const char *outdir, int nfiles)
{
    if (nfiles == 1 && (!outdir || !*outdir)) return strdup(outfilenm ? outfilenm : "");
    char *stem = path_stem(in);
    if (!stem) return NULL;
    char *name;
    if (outdir && *outdir) {
        const char *bn = strrchr(stem, '/');
        bn = bn ? bn + 1 : stem;
        const size_t n = strlen(outdir) + strlen(bn) + 2;
        name = (char *)malloc(n);
        snprintf(name, n, "%s/%s", outdir, bn);
        free(stem);
    } else {
        name = stem;
    }
    const size_t n2 = strlen(name) + 8;
    char *out = (char *)malloc(n2);
    snprintf(out, n2, "%s.cohout", name);
    free(name);
    return out;
}

static int cmp_str(const void *a, const void *b)
{
    return strcmp(*(char *const *)a, *(char *const *)b);
}

static int collisions(char **files, int n, const char *outdir)
{
    if (n <= 1) return 0;
    char **o = (char **)malloc((size_t)n * sizeof(char *));
    for (int i = 0; i < n; i++) o[i] = output_path(files[i], "", outdir, 2);
    qsort(o, (size_t)n, sizeof(char *), cmp_str);
    int bad = 0;
    for (int i = 1; i < n && !bad; i++)
        if (strcmp(o[i], o[i - 1]) == 0) {
            fprintf(stderr, "Error: several inputs would write the same candidate file "
                            "'%s'\n", o[i]);
            bad = 1;
        }
    for (int i = 0; i < n; i++) free(o[i]);
    free(o);
    return bad;
}

/* A .dat.list (src/cli.c:514-554 interface): one input per line, blank
 * lines and '#' comments skipped, space, tab, CR and LF trimmed at both ends;
 * a relative entry is taken relative to the list file's directory. */
struct list_acc {
    const char *listpath;
    int dirlen;           /* length of the list's directory, -1 for none */
    char **v;
    int n, cap;
};

static int list_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

static int list_line(const char *s, size_t len, void *ctx)
{
    struct list_acc *L = (struct list_acc *)ctx;
    while (len && list_blank(*s)) { s++; len--; }
    while (len && list_blank(s[len - 1])) len--;
    if (!len || *s == '#') return 0;
    char *full = NULL;
    if (*s == '/' || L->dirlen < 0) full = strndup(s, len);
    else if (asprintf(&full, "%.*s/%.*s", L->dirlen, L->listpath, (int)len, s) < 0) full = NULL;
    if (!full) return -1;
    if (L->n == L->cap) {
        const int cap = L->cap ? 2 * L->cap : 64;
        char **g = (char **)realloc(L->v, (size_t)cap * sizeof(char *));
        if (!g) { free(full); return -1; }
        L->v = g;
        L->cap = cap;
    }
    L->v[L->n++] = full;
    return 0;
}

static char **read_list(const char *listpath, int *nout)
{
    const char *sl = strrchr(listpath, '/');
    struct list_acc L = {listpath, sl && sl > listpath ? (int)(sl - listpath) : -1, NULL, 0, 0};
    if (cs_for_each_line(listpath, list_line, &L) != 0) {
        fprintf(stderr, "Error: cannot open list file '%s'\n", listpath);
        for (int i = 0; i < L.n; i++) free(L.v[i]);
        free(L.v);
        return NULL;
    }
    *nout = L.n;
    return L.v;
}

/* Every directory of path, outermost first (existing ones are left alone). */
static void mkdir_p(const char *path)
{
    if (!path || !*path) return;
    for (const char *sep = strchr(path + 1, '/'); sep; sep = strchr(sep + 1, '/')) {
        char *prefix = strndup(path, (size_t)(sep - path));
        if (!prefix) return;
        mkdir(prefix, 0777);
        free(prefix);
    }
    mkdir(path, 0777);
}


/* ---- input prefetch (src/cli.c:556-686) ---------------------------------
 * PF_PROD producers prepare input i -- the .fft mapped, or the .dat read and
 * realfft + rednoise'd -- into slot i % nslots as soon as the searcher has
 * released input i - nslots; the searcher takes the inputs in order.  Each
 * producer's transforms get an equal share of the OpenMP threads. */
enum { PF_SLOTS = 4, PF_PROD = 2 };

typedef struct Prefetch Prefetch;
typedef struct {
    Prefetch *pf;
    int first;
} PfArg;

struct Prefetch {
    char **paths;
    int n, nslots, nprod, nstarted;
    int released;                 /* inputs the searcher is done with */
    int stop;
    DatPipe *pipe[PF_SLOTS];
    FFTFile ft[PF_SLOTS];
    int held[PF_SLOTS];           /* index of the input in the slot, -1 for none */
    int ok[PF_SLOTS];
    pthread_mutex_t mu;
    pthread_cond_t cv;
    pthread_t th[PF_PROD];
    PfArg arg[PF_PROD];
};

static void *pf_producer(void *v)
{
    const PfArg *a = (const PfArg *)v;
    Prefetch *pf = a->pf;
    for (int i = a->first; i < pf->n; i += pf->nprod) {
        const int slot = i % pf->nslots;
        pthread_mutex_lock(&pf->mu);
        while (!pf->stop && i - pf->released >= pf->nslots) pthread_cond_wait(&pf->cv, &pf->mu);
        const int stop = pf->stop;
        pthread_mutex_unlock(&pf->mu);
        if (stop) break;
        int is_dat = 0;
        const int ok = cs_open_input(pf->paths[i], pf->pipe[slot], &pf->ft[slot], &is_dat) == 0;
        pthread_mutex_lock(&pf->mu);
        pf->ok[slot] = ok;
        pf->held[slot] = i;
        pthread_cond_broadcast(&pf->cv);
        pthread_mutex_unlock(&pf->mu);
    }
    return NULL;
}

static void pf_finish(Prefetch *pf)
{
    pthread_mutex_lock(&pf->mu);
    pf->stop = 1;
    pthread_cond_broadcast(&pf->cv);
    pthread_mutex_unlock(&pf->mu);
    for (int t = 0; t < pf->nstarted; t++) pthread_join(pf->th[t], NULL);
    pthread_mutex_destroy(&pf->mu);
    pthread_cond_destroy(&pf->cv);
    for (int s = 0; s < pf->nslots; s++) datpipe_free(pf->pipe[s]);
}

static int pf_start(Prefetch *pf, char **paths, int n, int device)
{
    memset(pf, 0, sizeof(*pf));
    pf->paths = paths;
    pf->n = n;
    pf->nslots = n < PF_SLOTS ? n : PF_SLOTS;
    pf->nprod = n < PF_PROD ? n : PF_PROD;
    pthread_mutex_init(&pf->mu, NULL);
    pthread_cond_init(&pf->cv, NULL);
    const int share = omp_get_max_threads() / pf->nprod;
    for (int s = 0; s < pf->nslots; s++) {
        pf->held[s] = -1;
        pf->pipe[s] = datpipe_create(6, 100, 6.0);
        if (!pf->pipe[s]) {
            pf_finish(pf);
            return -1;
        }
        datpipe_set_threads(pf->pipe[s], share > 1 ? share : 1);
        datpipe_set_device(pf->pipe[s], device);
    }
    for (int t = 0; t < pf->nprod; t++) {
        pf->arg[t].pf = pf;
        pf->arg[t].first = t;
        if (pthread_create(&pf->th[t], NULL, pf_producer, &pf->arg[t]) != 0) {
            pf_finish(pf);
            return -1;
        }
        pf->nstarted++;
    }
    return 0;
}

/* input i, once prepared: 1 with *ft filled, 0 if it could not be opened */
static int pf_get(Prefetch *pf, int i, FFTFile *ft)
{
    const int slot = i % pf->nslots;
    pthread_mutex_lock(&pf->mu);
    while (pf->held[slot] != i) pthread_cond_wait(&pf->cv, &pf->mu);
    *ft = pf->ft[slot];
    const int ok = pf->ok[slot];
    pthread_mutex_unlock(&pf->mu);
    return ok;
}

static void pf_release(Prefetch *pf)
{
    pthread_mutex_lock(&pf->mu);
    pf->released++;
    pthread_cond_broadcast(&pf->cv);
    pthread_mutex_unlock(&pf->mu);
}

static int run(int argc, char **argv, int list)
{
    Opts o;
    memset(&o, 0, sizeof(o));
    o.threshold = 8.0; o.outfilenm = ""; o.outdir = ""; o.zaplist = ""; o.nharms = 60;
    o.ncands = 100;
    o.lobin = 100; o.lofreq = 0.1; o.hifreq = 125.0; o.hidr = 0.5; o.m = 16;
    o.precision = CS_PREC_F32; o.sigma = CS_SIGMA_ANALYTIC; o.maxdecim = 6; o.drtol = 1.0;
    o.numharm = 16; o.remove = 1; o.harm_remove = 1;
    o.files = (char **)calloc((size_t)argc + 1, sizeof(char *));
    for (int i = 1; i < argc; i++) {
        char *arg = argv[i];
        if (!strcmp(arg, "--help") || !strcmp(arg, "-h")) { usage(stdout, list); return 0; }
        if (arg[0] != '-' || !strcmp(arg, "-")) { o.files[o.nfiles++] = arg; continue; }
        char *name = arg, *val = NULL;
        if (!strncmp(arg, "--", 2)) {
            char *eq = strchr(arg, '=');
            if (eq) { *eq = '\0'; val = eq + 1; }
        }
#define NEED() do { if (!val) { if (i + 1 >= argc) { fprintf(stderr, "Error: %s needs a value\n", name); return 2; } val = argv[++i]; } } while (0)
#define BAD(what) do { fprintf(stderr, "Error: bad %s '%s'\n", what, val); return 2; } while (0)
        if (!strcmp(name, "-t") || !strcmp(name, "--threshold")) { NEED(); if (parse_d(val, &o.threshold)) BAD("threshold"); }
        else if (!strcmp(name, "-o") || !strcmp(name, "--outputfilenm")) { NEED(); o.outfilenm = val; }
        else if (!strcmp(name, "--outdir")) { NEED(); o.outdir = val; }
        else if (!strcmp(name, "-n") || !strcmp(name, "--nharms")) { NEED(); if (parse_i(val, &o.nharms)) BAD("nharms"); }
        else if (!strcmp(name, "--ncands")) { NEED(); if (parse_i(val, &o.ncands)) BAD("ncands"); }
        else if (!strcmp(name, "--lobin")) { NEED(); if (parse_i(val, &o.lobin)) BAD("lobin"); }
        else if (!strcmp(name, "--lofreq")) { NEED(); if (parse_d(val, &o.lofreq)) BAD("lofreq"); }
        else if (!strcmp(name, "--hifreq")) { NEED(); if (parse_d(val, &o.hifreq)) BAD("hifreq"); }
        else if (!strcmp(name, "--hidr")) { NEED(); if (parse_d(val, &o.hidr)) BAD("hidr"); }
        else if (!strcmp(name, "--m")) { NEED(); if (parse_i(val, &o.m)) BAD("m"); }
        else if (!strcmp(name, "--precision")) {
            NEED();
            if (!strcmp(val, "f32")) o.precision = CS_PREC_F32;
            else if (!strcmp(val, "f64")) o.precision = CS_PREC_F64;
            else { fprintf(stderr, "Error: precision must be f32 or f64\n"); return 2; }
        } else if (!strcmp(name, "--sigma")) {
            NEED();
            if (!strcmp(val, "analytic")) o.sigma = CS_SIGMA_ANALYTIC;
            else if (!strcmp(val, "measured")) o.sigma = CS_SIGMA_MEASURED;
            else { fprintf(stderr, "Error: sigma must be analytic or measured\n"); return 2; }
        } else if (!strcmp(name, "--blocksize")) { NEED(); if (parse_i(val, &o.blocksize)) BAD("blocksize"); }
        else if (!strcmp(name, "--maxdecim")) { NEED(); if (parse_i(val, &o.maxdecim)) BAD("maxdecim"); }
        else if (!strcmp(name, "--drtol")) { NEED(); if (parse_d(val, &o.drtol)) BAD("drtol"); }
        else if (!strcmp(name, "--numharm")) { NEED(); if (parse_i(val, &o.numharm)) BAD("numharm"); }
        else if (!strcmp(name, "--zaplist")) { NEED(); o.zaplist = val; }
        else if (!strcmp(name, "--baryv")) { NEED(); if (parse_d(val, &o.baryv)) BAD("baryv"); }
        else if (!strcmp(name, "--threads")) {
            NEED();
            int t = 0;
            if (parse_i(val, &t) || t < 1) { fprintf(stderr, "Error: --threads expects a positive integer\n"); return 2; }
        } else if (!strcmp(name, "--wisdomfile")) { NEED(); }
        else if (!strcmp(name, "--noremove")) o.remove = 0;
        else if (!strcmp(name, "--noharmremove")) o.harm_remove = 0;
        else if (!strcmp(name, "-v") || !strcmp(name, "--verbose")) o.verbose = 1;
        else if (!strcmp(name, "--progressbar") || !strcmp(name, "--noprogress") ||
                 !strcmp(name, "--nowisdom") || !strcmp(name, "--noplot")) { }
        else if (!strcmp(name, "--gpu")) {
            if (val) {
                if (parse_i(val, &o.gpu_device) || o.gpu_device < 0) {
                    fprintf(stderr, "Error: --gpu expects a non-negative device index\n");
                    return 2;
                }
            } else if (i + 1 < argc) {
                char *e = NULL;
                const long v = strtol(argv[i + 1], &e, 10);
                if (e && e != argv[i + 1] && *e == '\0' && v >= 0) { o.gpu_device = (int)v; i++; }
            }
        } else if (!strcmp(name, "--host-staging")) o.host_staging = 1;
        else if (!strcmp(name, "--metricstats")) o.metricstats = 1;
        else if (!strcmp(name, "--normalize")) o.normalize = 1;
        else if (!strcmp(name, "--plot") || !strcmp(name, "--plotstem") ||
                   !strcmp(name, "--plotcols") || !strcmp(name, "--plotrows") ||
                   !strcmp(name, "--plot-candidates") || !strcmp(name, "--prime-wisdom")) {
            fprintf(stderr, "Error: option '%s' is not supported by this GPU build\n", name);
            return 2;
        } else {
            fprintf(stderr, "Error: unknown option '%s'\n", name);
            usage(stderr, list);
            return 2;
        }
#undef NEED
#undef BAD
    }
    if (list) {
        if (o.nfiles != 1) { fprintf(stderr, "Error: coherent_search_list expects exactly one .list file\n"); return 2; }
        if (o.outfilenm[0]) { fprintf(stderr, "Error: -o is not allowed with a list (use --outdir)\n"); return 2; }
        int n = 0;
        char **v = read_list(o.files[0], &n);
        if (!v || n == 0) { fprintf(stderr, "Error: list file '%s' contains no inputs\n", o.files[0]); return 2; }
        free(o.files);
        o.files = v;
        o.nfiles = n;
        o.files_owned = 1;
    } else if (o.nfiles == 0) {
        fprintf(stderr, "Error: at least one FFT file is required\n");
        usage(stderr, list);
        return 2;
    }
    if (o.nfiles > 1 && o.outfilenm[0]) {
        fprintf(stderr, "Error: -o/--outputfilenm names one output file but %d FFT files were "
                        "given; omit it or use --outdir\n", o.nfiles);
        return 2;
    }
    if (o.m <= 0 || (o.m & 1)) { fprintf(stderr, "Error: m must be a positive even integer\n"); return 2; }
    if (o.nharms < 1) { fprintf(stderr, "Error: nharms must be positive\n"); return 2; }
    char gerr[256] = {0};
    if (cs_gpu_available(o.gpu_device, gerr, sizeof(gerr)) != 1) {
        fprintf(stderr, "Error: %s\n", gerr);
        return 2;
    }
    if (o.blocksize == 0) o.blocksize = CS_GPU_DEFAULT_BLOCKSIZE;
    if (collisions(o.files, o.nfiles, o.outdir)) return 2;
    if (o.outdir[0]) mkdir_p(o.outdir);

    SearchParams params;
    params_defaults(&params);
    params.nharms = o.nharms;
    params.m = o.m;
    params.hidr = o.hidr;
    params.threshold = o.threshold;
    params.precision = o.precision;
    params.sigma = o.sigma;
    params.ndecim = decimation_set(o.nharms, o.maxdecim, params.decimations, 64);

    ZapList zap;
    memset(&zap, 0, sizeof(zap));
    if (o.zaplist[0] && zaplist_read(o.zaplist, &zap) != 0) {
        fprintf(stderr, "Error: cannot read zaplist '%s'\n", o.zaplist);
        if (o.files_owned)
            for (int i = 0; i < o.nfiles; i++) free(o.files[i]);
        free(o.files);
        return 2;
    }
    SearchCache cache;
    memset(&cache, 0, sizeof(cache));
    Prefetch pf;
    int use_pf = 0;
    if (o.nfiles > 1) {
        use_pf = pf_start(&pf, o.files, o.nfiles, o.host_staging ? -1 : o.gpu_device) == 0;
        if (!use_pf) fprintf(stderr, "Warning: cannot start the prefetch threads; running sequentially\n");
    }
    DatPipe *pipe = use_pf ? NULL : datpipe_create(6, 100, 6.0);
    if (pipe) datpipe_set_device(pipe, o.host_staging ? -1 : o.gpu_device);
    int status = 0;
    for (int fi = 0; fi < o.nfiles; fi++) {
        const char *path = o.files[fi];
        FFTFile ft;
        int is_dat = 0;
        if (use_pf) {
            if (!pf_get(&pf, fi, &ft)) {
                fprintf(stderr, "Error: failed to prepare '%s'\n", path);
                status = 1;
                pf_release(&pf);
                continue;
            }
        } else if (cs_open_input(path, pipe, &ft, &is_dat) != 0) {
            fprintf(stderr, "Error: failed to open '%s'\n", path);
            status = 1;
            continue;
        }
        /* src/cli.c:1015-1024: the zaplist edits the in-memory spectrum */
        if (zap.n > 0 && zaplist_apply(&zap, &ft, o.baryv) < 0) {
            fprintf(stderr, "Error: failed to apply zaplist to '%s'\n", path);
            status = 1;
            fftfile_close(&ft);
            if (use_pf) pf_release(&pf);
            continue;
        }
        /* src/cli.c:1049-1058 */
        const int use_ms = o.metricstats || o.normalize;
        MetricStats ms;
        if (use_ms) metricstats_init(&ms, 0.0, 64.0, 3200, 16);
        CandidateList c = search_file(&ft, &params, o.lofreq, o.hifreq, o.lobin, o.blocksize,
                                      o.threshold, o.remove, o.drtol, o.harm_remove,
                                      o.numharm, use_ms ? &ms : NULL, o.normalize, o.verbose, 0,
                                      NULL, &cache, 1.0, o.gpu_device);
        candlist_sort_metric_desc(&c);
        if ((int)c.n > o.ncands) c.n = (size_t)o.ncands;
        CandidateList d = measure_ducy(&ft, &c, &params);
        candlist_free(&c);
        const int names = list ? (o.nfiles > 1 ? o.nfiles : 2) : o.nfiles;
        char *outfile = output_path(path, o.outfilenm, o.outdir, names);
        if (write_candidates(&d, outfile ? outfile : "", o.threshold) != 0) status = 1;
        if (use_ms) {
            /* report base: -o, else the .cohout path without its suffix, else the
             * input's name without directory and extension (src/cli.c:1075-1091) */
            char *base = NULL;
            if (o.outfilenm[0]) {
                base = strdup(o.outfilenm);
            } else if (outfile && *outfile) {
                base = strdup(outfile);
                char *dot = base ? strrchr(base, '.') : NULL;
                if (dot && strcmp(dot, ".cohout") == 0) *dot = '\0';
            } else {
                char *stem = path_stem(ft.path ? ft.path : path);
                if (stem) {
                    const char *bn = strrchr(stem, '/');
                    base = strdup(bn ? bn + 1 : stem);
                    free(stem);
                }
            }
            report_metricstats(&ms, base ? base : "candidates", &params, o.threshold);
            free(base);
            metricstats_free(&ms);
        }
        free(outfile);
        candlist_free(&d);
        fftfile_close(&ft);
        if (use_pf) pf_release(&pf);
    }
    if (use_pf) pf_finish(&pf);
    search_cache_free(&cache);
    datpipe_free(pipe);
    zaplist_free(&zap);
    if (o.files_owned)
        for (int i = 0; i < o.nfiles; i++) free(o.files[i]);
    free(o.files);
    return status;
}

int cs_main(int argc, char **argv) { return run(argc, argv, 0); }
int cs_main_list(int argc, char **argv) { return run(argc, argv, 1); }
