/*
 * PRESTO .inf / .fft input (the reference's FFTFile interface,
 * include/coherent_search.h:48-81) and small writers for synthetic inputs.
 *
 * Interface kept from src/fileio.c:38-154 (it is what a caller observes):
 *   - an .inf field is recognised when the raw line starts with its key,
 *     leading space included; its value is the text after the line's last
 *     '=' (the whole line without one), surrounding white space removed; a
 *     later line of the same key wins;
 *   - name.fft pairs with the name whose last four characters are ".inf";
 *   - the .fft is mapped copy-on-write (MAP_PRIVATE), so a zaplist edits the
 *     spectrum in memory only; namps = bytes / 8, T = N dt, df = 1/T, DC and
 *     Nyquist are the first two floats, "_red.fft" marks a dereddened file;
 *   - the error messages and their order of checks.
 *
 * The reader itself is this file's own: the .inf is read whole into memory
 * and scanned once against a table of fields, and the mapping of the
 * spectrum is a helper of its own.
 */
#define _GNU_SOURCE
#include "host_internal.h"

#include <ctype.h>
#include <errno.h>
#include <fcntl.h>
#include <stddef.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

static char *copy_span(const char *s, size_t n)
{
    char *r = (char *)malloc(n + 1);
    if (!r) return NULL;
    memcpy(r, s, n);
    r[n] = '\0';
    return r;
}

/* the whole file as a NUL-terminated buffer (its length in *len) */
static char *slurp(const char *path, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    size_t cap = 4096, n = 0;
    char *buf = (char *)malloc(cap);
    while (buf) {
        n += fread(buf + n, 1, cap - n - 1, f);
        if (n < cap - 1) break;
        char *g = (char *)realloc(buf, cap * 2);
        if (!g) {
            free(buf);
            buf = NULL;
            break;
        }
        buf = g;
        cap *= 2;
    }
    const int bad = ferror(f);
    fclose(f);
    if (!buf || bad) {
        free(buf);
        return NULL;
    }
    buf[n] = '\0';
    *len = n;
    return buf;
}

int cs_for_each_line(const char *path, int (*fn)(const char *, size_t, void *), void *ctx)
{
    size_t n = 0;
    char *text = slurp(path, &n);
    if (!text) return -1;
    int rc = 0;
    for (const char *s = text, *end = text + n; s < end && rc == 0;) {
        const char *nl = memchr(s, '\n', (size_t)(end - s));
        const char *e = nl ? nl : end;
        rc = fn(s, (size_t)(e - s), ctx) != 0 ? -1 : 0;
        s = nl ? nl + 1 : end;
    }
    free(text);
    return rc;
}

enum inf_kind { INF_TEXT, INF_REAL, INF_COUNT };

/* The fields the search reads, by the key their line starts with. */
static const struct {
    const char *key;
    enum inf_kind kind;
    size_t value;         /* offset of the value in SimpleInf */
    ptrdiff_t seen;       /* offset of its has_ flag, -1 for none */
} inf_fields[] = {
    {" Object being observed", INF_TEXT, offsetof(SimpleInf, object), -1},
    {" Epoch", INF_REAL, offsetof(SimpleInf, epoch), offsetof(SimpleInf, has_epoch)},
    {" Number of bins", INF_COUNT, offsetof(SimpleInf, N), offsetof(SimpleInf, has_N)},
    {" Width of each time series bin", INF_REAL, offsetof(SimpleInf, dt),
     offsetof(SimpleInf, has_dt)},
    {" Dispersion measure", INF_REAL, offsetof(SimpleInf, DM), offsetof(SimpleInf, has_DM)},
};

/* one line [s, e): its value span, after the last '=' and trimmed */
static void line_value(const char *s, const char *e, const char **vs, const char **ve)
{
    const char *v = s;
    for (const char *c = s; c < e; c++)
        if (*c == '=') v = c + 1;
    while (v < e && isspace((unsigned char)*v)) v++;
    while (e > v && isspace((unsigned char)e[-1])) e--;
    *vs = v;
    *ve = e;
}

static void inf_take(SimpleInf *inf, const char *s, const char *e)
{
    const size_t len = (size_t)(e - s);
    for (size_t k = 0; k < sizeof(inf_fields) / sizeof(inf_fields[0]); k++) {
        const size_t kl = strlen(inf_fields[k].key);
        if (len < kl || memcmp(s, inf_fields[k].key, kl) != 0) continue;
        const char *vs, *ve;
        line_value(s, e, &vs, &ve);
        char *val = copy_span(vs, (size_t)(ve - vs));
        if (!val) return;
        char *field = (char *)inf + inf_fields[k].value;
        switch (inf_fields[k].kind) {
        case INF_TEXT:
            free(*(char **)field);
            *(char **)field = val;
            val = NULL;
            break;
        case INF_REAL:
            *(double *)field = strtod(val, NULL);
            break;
        case INF_COUNT:
            *(long long *)field = strtoll(val, NULL, 10);
            break;
        }
        free(val);
        if (inf_fields[k].seen >= 0) *(int *)((char *)inf + inf_fields[k].seen) = 1;
        return;    /* the first key that matches owns the line */
    }
}

int inf_read(const char *path, SimpleInf *inf)
{
    memset(inf, 0, sizeof(*inf));
    inf->path = path ? copy_span(path, strlen(path)) : NULL;
    size_t n = 0;
    char *text = path ? slurp(path, &n) : NULL;
    if (!text) return -1;
    const char *end = text + n;
    for (const char *s = text; s < end;) {
        const char *nl = memchr(s, '\n', (size_t)(end - s));
        const char *e = nl ? nl + 1 : end;     /* the line keeps its newline, as read */
        inf_take(inf, s, e);
        s = e;
    }
    free(text);
    return 0;
}

void inf_free(SimpleInf *inf)
{
    if (!inf) return;
    free(inf->path);
    free(inf->object);
    memset(inf, 0, sizeof(*inf));
}

/* Map a file copy-on-write.  0 with *map / *bytes set (NULL, 0 for an empty
 * file); -1 after printing the failing call's message. */
static int map_private(const char *path, void **map, size_t *bytes)
{
    *map = NULL;
    *bytes = 0;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        fprintf(stderr, "Error: cannot open '%s': %s\n", path, strerror(errno));
        return -1;
    }
    struct stat sb;
    int rc = 0;
    if (fstat(fd, &sb) != 0) {
        fprintf(stderr, "Error: cannot stat '%s': %s\n", path, strerror(errno));
        rc = -1;
    } else if (sb.st_size > 0) {
        void *m = mmap(NULL, (size_t)sb.st_size, PROT_READ | PROT_WRITE, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            fprintf(stderr, "Error: cannot mmap '%s': %s\n", path, strerror(errno));
            rc = -1;
        } else {
            *map = m;
        }
    }
    if (rc == 0) *bytes = (size_t)sb.st_size;
    close(fd);
    return rc;
}

int fftfile_open(FFTFile *ft, const char *path)
{
    memset(ft, 0, sizeof(*ft));
    ft->path = copy_span(path, strlen(path));
    /* the companion .inf: the path's last four characters become ".inf" */
    const int stem = (int)(strlen(path) > 4 ? strlen(path) - 4 : 0);
    char *infpath = NULL;
    if (asprintf(&infpath, "%.*s.inf", stem, path) < 0) return -1;
    const int found = inf_read(infpath, &ft->inf) == 0;
    if (!found) fprintf(stderr, "Error: The .inf file '%s' was not found.\n", infpath);
    free(infpath);
    if (!found) return -1;
    const char *missing = !ft->inf.has_N ? "Number of bins"
                          : !ft->inf.has_dt ? "Width of each time series bin" : NULL;
    if (missing) {
        fprintf(stderr, "Error: Missing '%s' in %s\n", missing, ft->inf.path);
        return -1;
    }
    void *map;
    size_t bytes;
    if (map_private(path, &map, &bytes) != 0) return -1;
    ft->map = map;
    ft->mapsize = bytes;
    ft->amps = (float *)map;
    ft->namps = (long long)(bytes / (2 * sizeof(float)));
    ft->N = ft->inf.N;
    ft->T = ft->inf.dt * (double)ft->N;
    ft->df = 1.0 / ft->T;
    ft->dereddened = ft->detrended = strstr(path, "_red.fft") != NULL;
    if (ft->namps > 0) {
        ft->DC = ft->amps[0];
        ft->Nyquist = ft->amps[1];
    }
    return 0;
}

void fftfile_close(FFTFile *ft)
{
    if (!ft) return;
    if (ft->map && ft->mapsize) munmap(ft->map, ft->mapsize);
    else if (ft->owns_amps) free(ft->amps);
    inf_free(&ft->inf);
    free(ft->path);
    memset(ft, 0, sizeof(*ft));
}

/* Minimal PRESTO-style .inf with the fields the readers use. */
int cs_write_inf(const char *path, long long N, double dt, double DM, const char *object)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, " Data file name without suffix          =  %s\n", object ? object : "synthetic");
    fprintf(f, " Telescope used                         =  None (Artificial Data Set)\n");
    fprintf(f, " Object being observed                  =  %s\n", object ? object : "synthetic");
    fprintf(f, " Epoch of observation (MJD)             =  60000.000000000000\n");
    fprintf(f, " Barycentered?           (1=yes, 0=no)  =  1\n");
    fprintf(f, " Number of bins in the time series      =  %lld\n", N);
    fprintf(f, " Width of each time series bin (sec)    =  %.17g\n", dt);
    fprintf(f, " Any breaks in the data? (1=yes, 0=no)  =  0\n");
    fprintf(f, " Dispersion measure (cm-3 pc)           =  %.6f\n", DM);
    fprintf(f, " Type of observation (EM band)          =  Radio\n");
    fclose(f);
    return 0;
}

int cs_write_floats(const char *path, const float *v, long long n)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const size_t w = fwrite(v, sizeof(float), (size_t)n, f);
    fclose(f);
    return w == (size_t)n ? 0 : -1;
}
