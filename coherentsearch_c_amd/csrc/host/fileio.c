/*
 * PRESTO .inf / .fft readers (src/fileio.c:38-164 semantics: the value after
 * the last '=', keys matched at line start including the leading space; the
 * .fft is mapped MAP_PRIVATE so a zaplist could edit it in memory; namps =
 * bytes / 8; T = N * dt) and small writers for synthetic inputs.
 */
#define _GNU_SOURCE
#include "host_internal.h"

#include <ctype.h>
#include <errno.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

static char *dupstr(const char *s)
{
    if (!s) return NULL;
    const size_t n = strlen(s) + 1;
    char *r = (char *)malloc(n);
    if (r) memcpy(r, s, n);
    return r;
}

static char *value_of(char *line)
{
    char *eq = strrchr(line, '=');
    char *s = eq ? eq + 1 : line;
    while (*s && isspace((unsigned char)*s)) s++;
    char *e = s + strlen(s);
    while (e > s && isspace((unsigned char)e[-1])) *--e = '\0';
    return s;
}

int inf_read(const char *path, SimpleInf *inf)
{
    memset(inf, 0, sizeof(*inf));
    inf->path = dupstr(path);
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char *line = NULL;
    size_t cap = 0;
    while (getline(&line, &cap, f) >= 0) {
        if (strncmp(line, " Object being observed", 22) == 0) {
            free(inf->object);
            inf->object = dupstr(value_of(line));
        } else if (strncmp(line, " Epoch", 6) == 0) {
            inf->epoch = strtod(value_of(line), NULL);
            inf->has_epoch = 1;
        } else if (strncmp(line, " Number of bins", 15) == 0) {
            inf->N = strtoll(value_of(line), NULL, 10);
            inf->has_N = 1;
        } else if (strncmp(line, " Width of each time series bin", 30) == 0) {
            inf->dt = strtod(value_of(line), NULL);
            inf->has_dt = 1;
        } else if (strncmp(line, " Dispersion measure", 19) == 0) {
            inf->DM = strtod(value_of(line), NULL);
            inf->has_DM = 1;
        }
    }
    free(line);
    fclose(f);
    return 0;
}

void inf_free(SimpleInf *inf)
{
    if (!inf) return;
    free(inf->path);
    free(inf->object);
    memset(inf, 0, sizeof(*inf));
}

/* name.fft -> name.inf (the last four characters replaced) */
static char *inf_path_for(const char *path)
{
    const size_t plen = strlen(path);
    const size_t keep = plen > 4 ? plen - 4 : 0;
    char *p = (char *)malloc(keep + 5);
    if (!p) return NULL;
    memcpy(p, path, keep);
    memcpy(p + keep, ".inf", 5);
    return p;
}

int fftfile_open(FFTFile *ft, const char *path)
{
    memset(ft, 0, sizeof(*ft));
    ft->path = dupstr(path);
    char *ip = inf_path_for(path);
    if (!ip) return -1;
    if (inf_read(ip, &ft->inf) != 0) {
        fprintf(stderr, "Error: The .inf file '%s' was not found.\n", ip);
        free(ip);
        return -1;
    }
    free(ip);
    if (!ft->inf.has_N) {
        fprintf(stderr, "Error: Missing 'Number of bins' in %s\n", ft->inf.path);
        return -1;
    }
    if (!ft->inf.has_dt) {
        fprintf(stderr, "Error: Missing 'Width of each time series bin' in %s\n",
                ft->inf.path);
        return -1;
    }
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        fprintf(stderr, "Error: cannot open '%s': %s\n", path, strerror(errno));
        return -1;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        fprintf(stderr, "Error: cannot stat '%s': %s\n", path, strerror(errno));
        close(fd);
        return -1;
    }
    const size_t bytes = (size_t)st.st_size;
    ft->mapsize = bytes;
    if (bytes > 0) {
        void *m = mmap(NULL, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            fprintf(stderr, "Error: cannot mmap '%s': %s\n", path, strerror(errno));
            close(fd);
            return -1;
        }
        ft->map = m;
        ft->amps = (float *)m;
        ft->namps = (long long)(bytes / sizeof(float) / 2);
    }
    close(fd);
    ft->N = ft->inf.N;
    ft->T = (double)ft->N * ft->inf.dt;
    ft->df = 1.0 / ft->T;
    ft->dereddened = strstr(path, "_red.fft") != NULL;
    ft->detrended = ft->dereddened;
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
