// Internal interface of the batched level sweeps (fvs_batch.hip,
// fvs_batch_fast.hip) shared with the AMR hierarchy (fvs_amr.hip).
#pragma once

#include "fvs_internal.hpp"
#include "fvs_physics.hpp"
#include <vector>

namespace fvs {
namespace batch {

constexpr int kG = 2;  // ghost layers per patch side

struct PatchInfo {
  int64_t offset;  // element offset of cell (-2, -2), component 0
  int32_t nx, ny;
  double dx;
};

struct BatchArgs {
  const PatchInfo* patches;
  const int4* tiles;  // (patch, start along the sweep, start across, growth
                      // across: rows / columns -w .. n + w - 1 are updated)
  int ntiles;
  const double* in;
  double* out;
  const double* dt;
  Gas<double> g;
  int limiter;
  unsigned long long* partials;
};

// Batched sweeps of the tolerance-checked arithmetic modes (FVS_FP64_RCP,
// FVS_FP32_FLUX), compiled with multiply-add contraction in
// fvs_batch_fast.hip.
void LaunchSweepFast(hipStream_t s, int dir, int shape, const BatchArgs& a, unsigned grid,
                     int precision);

// One coarse face of a coarse-fine interface (a flux-register entry): the
// fine patch's `ratio` faces along it and the coarse face itself.  (fi, fj)
// is the fine patch's cell right of / above the first fine face, (ci, cj)
// the coarse patch's cell right of / above the coarse face (in-patch
// indices, ghosts allowed); side 0: the coarse cell that receives the
// correction lies left of / below the face (the fine patch's low side),
// side 1: right of / above it.
struct RegFace {
  int32_t fp, fi, fj;
  int32_t cp, ci, cj;
  int32_t side, pad;
};

// Adds sign * (dx * dt) * F of each register entry's faces (all `ratio`
// fine faces in order, or the one coarse face) to reg (component-major,
// 4 x nfaces), F the MUSCL-Hancock + HLLEM flux the batched sweep of
// direction dir computes from `pool` with the level's time step *dt.
struct RegArgs {
  const RegFace* faces;
  int first, count, nfaces;
  const PatchInfo* patches;
  const double* pool;
  const double* dt;
  double* reg;
  Gas<double> g;
  int limiter;
  int ratio;     // faces per entry: the refinement ratio (fine) or 1 (coarse)
  int fine;      // 1: fine faces (fp, fi, fj), 0: coarse face (cp, ci, cj)
  double sign;
};
// Register contributions a0 and a1 (count 0: none) in one launch.
void LaunchRegFlux(hipStream_t s, int dir, const RegArgs& a0, const RegArgs& a1, int precision);
void LaunchRegFluxFast(hipStream_t s, int dir, const RegArgs& a0, const RegArgs& a1,
                       int precision);

} // namespace batch
} // namespace fvs

struct fvs_batch {
  fvs_handle* h;
  int64_t count;
  std::vector<fvs::batch::PatchInfo> host_patches;
  fvs::batch::PatchInfo* patches;
  int64_t pool_size;
  double* pool[2];
  int shape[2];                 // chosen tile shape per direction
  int4* tiles[2];               // tile list per direction
  int ntiles[2];
  unsigned grid[2];
  int4* tiles_grown[2];         // the same tiles over the interior grown by
  int ntiles_grown[2];          // kG rows / columns across the direction
  unsigned grid_grown[2];
  unsigned long long* partials;
  int64_t partials_cap;
  double* dts;                  // device: [0] dt of the sweeps, [1..2] CFL results
  hipStream_t cap_stream;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int cus;
  int64_t gen;                  // bumped whenever the tile lists change
};

namespace fvs {
namespace batch {
// One batched sweep of direction dir from pool src into pool 1 - src on
// stream s; the minimum stable dt of its faces goes to *dt_out (nullable).
int Sweep(fvs_batch* b, hipStream_t s, int dir, int src, double* dt_out);
// The same sweep without its reduction: the per-workgroup minima (keys, at
// most 8 per CU) go to `partials`.  grown: the sweep also updates the kG
// ghost rows (dir 0) / columns (dir 1) on each side across the direction,
// reading the ghost layers along it; its CFL minimum covers the interior
// faces only.
int SweepPartials(fvs_batch* b, hipStream_t s, int dir, int src, unsigned long long* partials,
                  bool grown = false);
// Slots of `cap` partial minima: clear = true sets every key to +inf (before
// sweeps whose grids may be smaller than cap), clear = false reduces slot q
// into out[q].
int LaunchSlotMin(hipStream_t s, unsigned long long* partials, int nslots, int cap, double* out,
                  bool clear);
} // namespace batch
} // namespace fvs
