/* Device staging (realfft + rednoise) shared between cs_stage.hip and the
 * session layer. */
#ifndef CS_STAGE_H
#define CS_STAGE_H

#include <hip/hip_runtime.h>

#include "cs_gpu.h"

/* longest rednoise block (the pipeline's endwidth) the device path holds */
#define CS_STAGE_MAXBLK 1024

extern "C" __attribute__((visibility("hidden"))) int cs_stage_enqueue(CsStage *g,
                                                                       const float *samples,
                                                                       float2 *d_out,
                                                                       hipStream_t st);

extern "C" __attribute__((visibility("hidden"))) long long cs_stage_N(const CsStage *g);

#endif
