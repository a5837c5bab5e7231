// Internal interface between the grader and the pair scorer (not part of the
// C ABI): lets the fused feature + scoring pass queue the scorer's forward
// pass behind the feature kernels without host waits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pdfgrade.h"

namespace pdf {

hipStream_t mlp_stream(const pdf_mlp *m);
int mlp_device(const pdf_mlp *m);
// Checks the feature matrix and sizes the workspace for `rows` rows at a time.
int mlp_prepare(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, uint64_t rows);
// Queues the forward pass of rows [r0, r1) on the scorer's stream.
// zero_stages: bit s says features [64 s, 64 s + 64) are zero in every row
// (the scorer skips their products).
int mlp_queue_forward(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, uint64_t r0, uint64_t r1, float *score_dev,
                      uint32_t zero_stages = 0);

}  // namespace pdf
