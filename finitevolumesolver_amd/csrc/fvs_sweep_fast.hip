// Kernels of the tolerance-checked arithmetic modes (refined-reciprocal FP64
// and FP32 flux) of the fused sweeps and the CFL pass, compiled with
// multiply-add contraction (-ffp-contract=fast; see the build recipe).  The
// IEEE FP64 kernels and the host entry points come from the plain build of
// fvs_sweep.hip.
#define FVS_SWEEP_FAST_TU 1
#include "fvs_sweep.hip"
