/*
 * The general search kernel: every configuration the specialised register
 * kernels (fast_kernel.hpp) cannot hold -- harmonic counts above 128 (or
 * above 64 with m > 16 or in double precision), m up to 64, and window spans
 * beyond their LDS tile (hidr > 0.5).  One trial per lane; the harmonic
 * stack and the c2r values live in a global work buffer, element-interleaved
 * across lanes so that a wave's accesses coalesce; the c2r runs a host-built
 * op list that follows the project c2r's recursion exactly (prime-factor
 * split as re-indexing, decimation in time as butterflies with their input
 * twiddles), so profiles and metrics are bit-identical to the oracle.
 */
#ifndef CS_GENERAL_H
#define CS_GENERAL_H

#include <hip/hip_runtime.h>

#include <vector>

#include "cs_device.h"

#define CS_GEN_MAXH 512           /* harmonics */
#define CS_GEN_MAXC2R CS_MAXSLOT  /* c2rs per search: the base and non-divisor decimations */

/* op kinds of a c2r plan; every op names its operands in the plan's int pool */
enum { CS_OP_PRE = 0, CS_OP_BF2 = 1, CS_OP_BF4 = 2, CS_OP_BFP = 3, CS_OP_COPY = 4 };

typedef struct {
    int kind;
    int R;              /* radix (BF*), or the c2r size H (PRE) */
    int pool;           /* offset of the operands in the pool */
    int pad;
} CsGenOp;

/* One c2r of a search.  Its profile (2H values, x[2s] = z_s.re,
 * x[2s+1] = z_s.im) ends in the value slots zout .. zout + H - 1. */
typedef struct {
    int H, XS;          /* size, stack stride (the decimation k; 1 for the base) */
    int op0, nop;       /* its ops */
    int zout;
    int pad;
} CsGenC2R;

struct CsGenPlan {
    std::vector<CsGenOp> ops;
    std::vector<int> pool;
    std::vector<double> twc, tws;       /* cos, sin of each twiddle the ops name */
    std::vector<CsGenC2R> c2r;
    int nslots = 0;                     /* value slots per lane */
    int nscratch = 0;                   /* slots past them for butterflies of primes above 13 */
};

/* The plan of a search's c2rs: c2r 0 of size nh (stride 1), then one per
 * size-stride pair (Hk[i], ks[i]).  0 on success. */
int cs_gen_build_plan(int nh, int nextra, const int *Hk, const int *ks, CsGenPlan *out);

/* Kernel arguments: the FastArgs fields the general kernel reads, plus the
 * plan and the work buffer. */
struct GenArgs {
    const float2 *amps;
    long long namps;
    const float *wtab;
    const int2 *hinfo;
    const int4 *tiles;
    int ntiles;
    int q, pnum, ws, grp;
    int m, mp;                  /* taps; weights of row r at wtab[r*ws + mp - m ...], A at + mp */
    int nh;
    const ChunkDesc *chunks;
    CandRec *recs;
    unsigned int *counter;
    unsigned int cap;
    double threshold;
    int nslot;
    int slot_k[CS_MAXSLOT];
    int slot_Hk[CS_MAXSLOT];
    int slot_c2r[CS_MAXSLOT];   /* c2r index of the slot's profile, -1 for a fold of the base */
    unsigned int wmask[CS_MAXSLOT];
    void *samp;
    void *mout;
    long long mstride;
    float2 *stack_out;
    long long dump_t0, dump_n;
    const CsGenOp *ops;
    const int *pool;
    const void *tw;             /* (cos, sin) pairs in the profile precision */
    int nc2r;
    CsGenC2R c2r[CS_GEN_MAXC2R];
    int nslots;                 /* the plan's value slots; scratch runs follow */
    int nvslots;                /* value slots per lane, scratch included */
    void *work;                 /* nvslots values, then nh + 1 stack entries, per lane */
    long long nlanes;           /* lanes of the persistent grid */
    int given;                  /* 1: the stacks of tiles tile0 .. tile0 + nslice - 1 are in
                                   the work buffer already (the stack pass of
                                   split_kernel.hpp), workgroup u takes tile tile0 + u */
    int tile0, nslice;
    int fslot;                  /* first of nh / 2 value slots a divisor decimation's fold of
                                   the base profile is written to before its boxcar */
};

/* work buffer bytes for nvslots value slots and nh harmonics over nlanes lanes */
static inline size_t cs_gen_work_bytes(int nvslots, int nh, long long nlanes, bool f64)
{
    return (size_t)nlanes * ((size_t)nvslots * (f64 ? 16 : 8) + (size_t)(nh + 1) * 8);
}

/* workgroups of the persistent grid (each loops over tiles) */
#define CS_GEN_GRID 2048

extern "C" __attribute__((visibility("hidden"))) hipError_t cs_gen_launch(const GenArgs *a,
                                                                          bool f64,
                                                                          hipStream_t st);

#endif
