// Batched-sweep kernels of the tolerance-checked arithmetic modes
// (refined-reciprocal FP64 and FP32 flux), compiled with multiply-add
// contraction (-ffp-contract=fast; see the build recipe).  The IEEE FP64
// kernels and the C entry points come from the plain build of fvs_batch.hip.
#define FVS_BATCH_FAST_TU 1
#include "fvs_batch.hip"
