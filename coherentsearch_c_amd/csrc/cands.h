/*
 * Candidate lists and collapse steps shared by the HIP library and the host
 * C library (src/candidate.c:11-161).  The csx_ names are internal; the host
 * library re-exports them under the reference's names and the HIP library
 * keeps them hidden.
 */
#ifndef CS_CANDS_H
#define CS_CANDS_H

#include <stddef.h>

#include "cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

void csx_candlist_init(CandidateList *l);
void csx_candlist_free(CandidateList *l);
void csx_candlist_push(CandidateList *l, Candidate c);
void csx_candlist_append(CandidateList *dst, const CandidateList *src);
/* stable: on ties the earlier element stays first */
void csx_candlist_sort_metric_desc(CandidateList *l);
void csx_candlist_sort_freq(CandidateList *l);
void csx_candlist_sort_r(CandidateList *l);
/* runs within dr_tol bins of their predecessor keep their strongest member;
 * result sorted by frequency (src/candidate.c:98-123) */
CandidateList csx_remove_duplicates(const CandidateList *in, double dr_tol);
int csx_harmonically_related(double r1, double r2, int numharm, double tol);
/* greedy by descending metric, dropping m/n-related candidates; result sorted
 * by frequency (src/candidate.c:138-161) */
CandidateList csx_remove_harmonics(const CandidateList *in, int numharm, double tol);

#ifdef __cplusplus
}
#endif
#endif
