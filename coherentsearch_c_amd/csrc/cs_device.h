/*
 * Device-side data layout shared by the host session code and the kernels.
 *
 * HBM layout of one search (all resident for the life of a session):
 *   amps    float2[namps]                 the spectrum, interleaved re/im
 *   wtab    float[nh*q*WS]                per (harmonic h, residue res) the m
 *                                         interpolation weights W, then A.re,
 *                                         A.im (WS = m + 4, 16-byte rows)
 *   hinfo   int2[nh]                      (rfloor0, first residue with carry)
 *   tw      float2[TW_SIZE]               twiddles, N(N-1)/2 + j for N <= TW_NMAX
 *   chunks  ChunkDesc[nchunks]            i0, n, filled mask (512 bits), per-slot nvalid,
 *                                         per-slot per-width 1/(sigma*sd_w)
 *   tiles   int4[ntiles]                  (chunk, first trial, count, 0)
 *   recs    CandRec[cap] + counter        above-threshold records
 */
#ifndef CS_DEVICE_H
#define CS_DEVICE_H

#include <stdint.h>

#define CS_TILE 256            /* trials per workgroup (one per lane) */
#define CS_MAXSLOT 17          /* base + up to 16 decimations */
#define CS_NSEQ 16             /* width candidates 1,2,3,4,6,9,13,19,28,42,63,94,141,211,316,474 */
#define CS_TW_NMAX 257
#define CS_SIGCAP 8192         /* measured sigma: samples per (chunk, decimation slot) */
#define CS_TW_SIZE (CS_TW_NMAX * (CS_TW_NMAX - 1) / 2)
/* LDS window slots (float2) per workgroup for nh harmonics of m taps: covers
 * sum_h (ceil(255*h*lodr) + m + 2) whenever hidr <= 0.5, rounded to 256. */
#define CS_WIN_CAP(nh, m) (((64 * ((nh) + 1) + (nh) * ((m) + 4)) + 255) / 256 * 256)
/* Weight-table row stride in floats: m weights, A.re, A.im, padded to an odd
 * number of 16-byte units, so that the same float4 of different rows falls in
 * different L1 banks when a wave gathers rows at scattered residues. */
#define CS_WSP(m) ((((m) + 5) / 4) % 2 ? ((m) + 5) / 4 * 4 : (((m) + 5) / 4 + 1) * 4)

typedef struct {
    long long i0;
    int n;
    int pad;
    unsigned long long filled[8];   /* bit (h-1) & 63 of word (h-1) >> 6: harmonic h fits */
    int nvalid[CS_MAXSLOT];         /* trials below the decimated Nyquist */
    int valid[CS_MAXSLOT];          /* 1/sigma > 0 */
    float invsw[CS_MAXSLOT][CS_NSEQ];     /* f32 path: (float)(invsigma*shape_w) */
    double invswd[CS_MAXSLOT][CS_NSEQ];   /* f64 path: invsigma*shape_w */
} ChunkDesc;

typedef struct {
    unsigned long long key;         /* chunk << 40 | slot << 32 | j */
    double metric;                  /* the f32 metric widened exactly, or the f64 metric */
} CandRec;

/* harmonic counts with a fused kernel (each for M = 16 and 32 padded taps,
 * decimations up to 16) */
#define CS_NH_LIST(X)                                                                  \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)   \
    X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28)     \
    X(29) X(30) X(31) X(32) X(33) X(34) X(35) X(36) X(37) X(38) X(39) X(40) X(41)     \
    X(42) X(43) X(44) X(45) X(46) X(47) X(48) X(49) X(50) X(51) X(52) X(53) X(54)     \
    X(55) X(56) X(57) X(58) X(59) X(60) X(61) X(62) X(63) X(64)
/* harmonic counts above 64 (up to 128, where the c2r twiddle table ends)
 * with a register profile pass of the two-pass search (f32 profiles; the
 * stack pass takes M = 16, 32 or 64 padded taps): the stack and profiles take
 * about 440 registers, one wave per SIMD */
#define CS_NH_BIG_LIST(X)                                                              \
    X(65) X(66) X(67) X(68) X(69) X(70) X(71) X(72) X(73) X(74) X(75) X(76) X(77)     \
    X(78) X(79) X(80) X(81) X(82) X(83) X(84) X(85) X(86) X(87) X(88) X(89) X(90)     \
    X(91) X(92) X(93) X(94) X(95) X(96) X(97) X(98) X(99) X(100) X(101) X(102)        \
    X(103) X(104) X(105) X(106) X(107) X(108) X(109) X(110) X(111) X(112) X(113)      \
    X(114) X(115) X(116) X(117) X(118) X(119) X(120) X(121) X(122) X(123) X(124)      \
    X(125) X(126) X(127) X(128)
#define CS_KMAX 16

static inline int cs_tw_off(int N) { return N * (N - 1) / 2; }

#endif
