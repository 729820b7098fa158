/*
 * fvs.h -- C ABI of the MI355X (gfx950) split patch update for the
 * perfect-gas Euler equations.
 *
 * Every entry point is stream-ordered on the HIP stream given to fvs_create,
 * takes plain pointers to device memory that the caller owns, never frees a
 * caller buffer and never throws.  Status codes mirror the reference's error
 * conventions:
 *
 *   FVS_OK                   success
 *   FVS_TIME_STEP_TOO_LARGE  CFL check failed; mirrors TimeStepTooLarge
 *                            (include/fub/TimeStepError.hpp:71-74)
 *   FVS_NEWTON_NOT_CONVERGED Newton iteration hit its iteration cap; mirrors
 *                            the throw in include/fub/NewtonIteration.hpp:44-50
 *   FVS_BAD_ARGUMENT         shape / enum / pointer error (nothing launched)
 *   FVS_HIP_ERROR            a HIP runtime call failed
 *
 * Layout.  States are component-major structure-of-arrays with x fastest, as
 * the reference's PatchDataView with layout_left over an AMReX FArrayBox
 * (include/fub/PatchDataView.hpp:184-356).  A Complete state of a D-dimensional
 * perfect gas has D+4 components (rho, rho*u_1..rho*u_D, rho*E, p, a), a
 * Conservative state D+2 (include/fub/equations/PerfectGas.hpp:132-161).
 * Directions are 0 = X, 1 = Y, 2 = Z.
 */
#ifndef FVS_H
#define FVS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fvs_status {
  FVS_OK = 0,
  FVS_TIME_STEP_TOO_LARGE = 1,
  FVS_NEWTON_NOT_CONVERGED = 2,
  FVS_BAD_ARGUMENT = 3,
  FVS_HIP_ERROR = 4
};

/* Base (two-state) numerical flux.
 * FVS_BASE_HLLEM_LARROUTUROU: perfect_gas::Hllem<Eq, true>, the default
 *   (include/fub/equations/perfect_gas/HllemMethod.hpp:30,92;
 *    src/equations/perfect_gas/HllemMethod.hpp:711-1072).
 * FVS_BASE_HLLEM: perfect_gas::Hllem<Eq, false>.
 * FVS_BASE_GODUNOV_EXACT: GodunovMethod<Eq, ExactRiemannSolver<Eq>>
 *   (include/fub/flux_method/GodunovMethod.hpp:130-199,
 *    src/equations/perfect_gas/ExactRiemannSolver.hpp:34-353).
 * FVS_BASE_HLL_EINFELDT: HllMethod<Eq, EinfeldtSignalVelocities<Eq>>
 *   (include/fub/flux_method/HllMethod.hpp:159-347,
 *    src/equations/perfect_gas/EinfeldtSignalVelocities.hpp:24-97); the
 *   patch operators (fvs_numeric_fluxes, fvs_stable_dt) only: the fused
 *   sweeps and the level drivers reject it. */
enum fvs_base_method {
  FVS_BASE_HLLEM_LARROUTUROU = 0,
  FVS_BASE_HLLEM = 1,
  FVS_BASE_GODUNOV_EXACT = 2,
  FVS_BASE_HLL_EINFELDT = 3
};

/* Reconstruction wrapped around the base flux.
 * FVS_RECON_NONE: FluxMethod<Base> (stencil width 1).
 * FVS_RECON_MUSCL_HANCOCK: MusclHancockMethod<Eq, Base, Limiter>
 *   (include/fub/flux_method/MusclHancockMethod.hpp:277-571, stencil width 2).
 * FVS_RECON_MUSCL_HANCOCK_PRIM: perfect_gas::MusclHancockPrim<Dim>
 *   (src/equations/perfect_gas/MusclHancockPrimMethod.hpp:59-176, stencil
 *   width 2): limited primitive slopes, characteristic half step, HLLEM
 *   (Larrouturou) on the two face states; its own limiter (the limiter field
 *   is ignored).  Requires FVS_BASE_HLLEM_LARROUTUROU and FVS_FP64; patch
 *   operators (fvs_numeric_fluxes, fvs_stable_dt) only.
 * FVS_RECON_MUSCL_HANCOCK_CHARACTERISTIC: perfect_gas::
 *   MusclHancockCharacteristic<Rank> with its default VanLeer limiter
 *   (src/equations/perfect_gas/MusclHancockCharacteristicMethod.hpp:26-320):
 *   limited wave amplitudes advanced along the characteristics over half a
 *   step, then HLLEM (Larrouturou).  Same restrictions as the PRIM form. */
enum fvs_reconstruction {
  FVS_RECON_NONE = 0,
  FVS_RECON_MUSCL_HANCOCK = 1,
  FVS_RECON_MUSCL_HANCOCK_PRIM = 2,
  FVS_RECON_MUSCL_HANCOCK_CHARACTERISTIC = 3
};

/* Slope limiters, include/fub/flux_method/MusclHancockMethod.hpp:31-187. */
enum fvs_limiter {
  FVS_LIMITER_MINMOD = 0,
  FVS_LIMITER_VANLEER = 1,
  FVS_LIMITER_NONE = 2, /* NoLimiter: central slope 0.5 (sL + sR) */
  FVS_LIMITER_NO_GRADIENT = 3
};

/* Arithmetic of reconstruction and flux.  The conservative update, Newton
 * iterations and cut-cell divisions are FP64 in every mode.
 * FVS_FP64:      IEEE double, correctly rounded division and square root (the
 *                reference's arithmetic; results are bit-identical to it).
 * FVS_FP32_FLUX: float reconstruction and flux with the hardware reciprocal
 *                and square-root estimates; FP64 CFL and accumulation.
 * FVS_FP64_RCP:  double throughout, but each division is a product with a
 *                reciprocal refined from the hardware estimate (shared by all
 *                quotients of one denominator) and each square root comes from
 *                a refined reciprocal square root: every result is within a
 *                few units in the last place, so a sweep agrees with FVS_FP64
 *                to about 1e-15 relative.  Fused sweeps, level drivers and
 *                batched sweeps honour it; the operator entry points compute
 *                as FVS_FP64. */
enum fvs_precision { FVS_FP64 = 0, FVS_FP32_FLUX = 1, FVS_FP64_RCP = 2 };

/* Newton initial guess of the exact Riemann solver. */
enum fvs_newton_guess {
  FVS_GUESS_MEAN_PRESSURE = 0, /* 0.5 (pL + pR), ExactRiemannSolver.hpp:91 */
  FVS_GUESS_LEARNED = 1        /* small MLP, see fvs_mlp_* below */
};

typedef struct fvs_method_desc {
  int dim;            /* 2 or 3 */
  double gamma;       /* ratio of specific heats, 1.4 in the reference */
  int base;           /* enum fvs_base_method */
  int reconstruction; /* enum fvs_reconstruction */
  int limiter;        /* enum fvs_limiter */
  int precision;      /* enum fvs_precision */
  int quirks;         /* 1: reproduce the reference's 3D HLLEM normal-momentum
                         anti-diffusion with roeU[iy] (HllemMethod.hpp:990-992);
                         0: use roeU[ix] as the 2D branch does */
  int newton_guess;   /* enum fvs_newton_guess */
  int max_newton_iterations; /* 5000 in the reference */
  double newton_tolerance;   /* 1e-7 in the reference (absolute |f|) */
  int boundary_solver;       /* enum fvs_boundary_solver (cut-cell methods) */
} fvs_method_desc;

/* Riemann solver of the cut-cell boundary problems, the RiemannSolver
 * template argument of KbnCutCellMethod
 * (include/fub/cutcell_method/KbnStabilisation.hpp:53-55).
 * FVS_BOUNDARY_EXACT: ExactRiemannSolver (the template default).
 * FVS_BOUNDARY_HLLEM: Hllem::SolveRiemannProblem with the handle's
 *   Larrouturou setting (src/equations/perfect_gas/HllemMethod.hpp:25-369),
 *   as examples/AMReX/EB/2D/Ramp.cpp:92-95 constructs it. */
enum fvs_boundary_solver { FVS_BOUNDARY_EXACT = 0, FVS_BOUNDARY_HLLEM = 1 };

/* Strided view over a box of cells or faces, the analogue of
 * PatchDataView<double, R+1, layout_stride>.  Element (i, j, k, c) lives at
 * data + (i-lo[0])*stride[0] + (j-lo[1])*stride[1] + (k-lo[2])*stride[2]
 * + c*comp_stride.  Boxes are inclusive [lo, hi]; unused dimensions have
 * lo = hi = 0. */
typedef struct fvs_view {
  double* data;
  int ncomp;
  int64_t lo[3];
  int64_t hi[3];
  int64_t stride[3];
  int64_t comp_stride;
} fvs_view;

/* Embedded-boundary geometry of one patch, CutCellData<Rank>
 * (include/fub/CutCellData.hpp:48-75), as views over cells or faces:
 *   volume_fractions     cells, 1 component (alpha)
 *   face_fractions[d]    faces normal to d, 1 component (beta)
 *   boundary_normals     cells, D components, AMReX orientation (pointing
 *                        into the body; GetBoundaryNormal negates it,
 *                        src/CutCellData.cpp:90-105)
 *   boundary_centroids   cells, D components, cell-relative in [-1/2, 1/2]
 *   unshielded[d], shielded_left[d], shielded_right[d], doubly_shielded[d]
 *                        faces normal to d, 2 components: 0 the absolute
 *                        fraction, 1 the fraction relative to beta, the
 *                        layout of src/AMReX/cutcell/PatchHierarchy.cpp:395-431
 * Views of unused directions may be empty (data NULL). */
typedef struct fvs_cutcell_data {
  fvs_view volume_fractions;
  fvs_view face_fractions[3];
  fvs_view boundary_normals;
  fvs_view boundary_centroids;
  fvs_view unshielded[3];
  fvs_view shielded_left[3];
  fvs_view shielded_right[3];
  fvs_view doubly_shielded[3];
} fvs_cutcell_data;

typedef struct fvs_handle fvs_handle;

/* Creates a flux-method handle bound to a HIP stream (NULL = default). */
int fvs_create(const fvs_method_desc* desc, void* hip_stream, fvs_handle** out);
int fvs_destroy(fvs_handle* handle);
/* Text of the last error on this thread (never NULL). */
const char* fvs_last_error(void);
/* Version string of the library build. */
const char* fvs_version(void);

/* ---- patch-level operators (reference API surface) -------------------- */

/* FluxMethod<Base>::ComputeStableDt(simd, states, dx, dir)
 * (include/fub/flux_method/FluxMethod.hpp:341-382): the minimum over all
 * faces of the states box (shrunk by the stencil in dir) of the base
 * method's stable dt.  The result is written to *dt_device (device memory). */
int fvs_stable_dt(fvs_handle* h, const fvs_view* states, double dx, int dir,
                  double* dt_device);

/* FluxMethod<Base>::ComputeNumericFluxes(simd, fluxes, states, dt, dx, dir)
 * (include/fub/flux_method/FluxMethod.hpp:246-293).  fluxes has D+2
 * components, states D+4; every face of fluxes' box is computed from the
 * stencil of 2*width cells around it. */
int fvs_numeric_fluxes(fvs_handle* h, const fvs_view* fluxes,
                       const fvs_view* states, double dt, double dx, int dir);

/* HyperbolicPatchIntegrator<Tag>::UpdateConservatively
 * (src/HyperbolicPatchIntegrator.cpp:35-154):
 * next = prev + dt/dx (F_left - F_right) over next's box; next and prev may
 * alias.  Components: the first ncomp of next. */
int fvs_update_conservatively(fvs_handle* h, const fvs_view* next,
                              const fvs_view* prev, const fvs_view* fluxes,
                              double dt, double dx, int dir);

/* CompleteFromConsFn (include/fub/CompleteFromCons.hpp:141-204) with
 * PerfectGas::CompleteFromCons (src/equations/PerfectGas.cpp:112-124):
 * copies the D+2 conservative components of cons into complete and
 * recomputes pressure and speed of sound.  complete and cons may alias. */
int fvs_complete_from_cons(fvs_handle* h, const fvs_view* complete,
                           const fvs_view* cons);

/* amrex::TransmissiveBoundary::FillBoundary
 * (src/AMReX/boundary_condition/TransmissiveBoundary.cpp:31-93): every cell
 * of the view's box that lies outside [domain_lo, domain_hi] in direction dir
 * gets the value of the domain's boundary cell of its row, all components. */
int fvs_transmissive_fill(fvs_handle* h, const fvs_view* states,
                          const int64_t domain_lo[3], const int64_t domain_hi[3],
                          int dir);

/* ---- fused single-patch level integrator -------------------------------- */

/* Splitting methods (src/split_method/GodunovSplitting.cpp:25-34,
 * src/split_method/StrangSplitting.cpp:25-69). */
enum fvs_splitting { FVS_SPLIT_GODUNOV = 0, FVS_SPLIT_STRANG = 1 };

/* A single patch of n[0] x n[1] x n[2] cells (n[2] = 1 in 2D) with
 * transmissive boundaries on every side, stored as D+2 conservative
 * component arrays without ghost cells (x fastest).  Pressure and speed of
 * sound are recomputed from the conservative variables wherever the
 * reference reads its Complete state; this is bit-identical because the
 * reference derives them with CompleteFromCons after every update. */
typedef struct fvs_patch_desc {
  int64_t n[3];
  double dx[3];
  int splitting; /* enum fvs_splitting */
  double cfl;    /* RunOptions::cfl, 0.8 in the reference */
  double smallest_dt; /* RunOptions::smallest_time_step_size
                         (include/fub/RunSimulation.hpp:61-63, 97, 109): a run
                         stops once t + smallest_dt >= final_time; 0 (or
                         less) selects the reference's default 1e-12 */
} fvs_patch_desc;

typedef struct fvs_level fvs_level;

/* Allocates the level's three state buffers and its device control block and
 * uploads nothing.  The state starts zeroed.  HLLEM (with MUSCL-Hancock or
 * first order) and first-order exact Godunov run the fused sweep kernels.
 * The other methods (MUSCL-Hancock over exact Godunov, the reference's
 * default MusclHancockMethod; the HLL base; MusclHancockPrim;
 * MusclHancockCharacteristic) get a level composed from the patch operators
 * above on one ghosted scratch of complete states, in the same sequence and
 * with the same results as the reference's level integrator; it synchronises
 * with the host a few times per step, ignores batch, speed hints, graphs and
 * timing, and has no contiguous state arrays (fvs_level_state returns NULL;
 * use fvs_level_get_state). */
int fvs_level_create(fvs_handle* h, const fvs_patch_desc* patch,
                     fvs_level** out);
int fvs_level_destroy(fvs_level* level);
/* Device pointer of component c of the current state (valid until the next
 * advance call); NULL for a composed level. */
double* fvs_level_state(fvs_level* level, int c);
/* Copies D+2 component arrays (device memory, ncell each) into the current
 * state; resets time and cycle count to (t0, 0). */
int fvs_level_set_state(fvs_level* level, const double* const* comps, double t0);
/* Copies the current state out into D+2 device arrays. */
int fvs_level_get_state(fvs_level* level, double* const* comps);

/* RunSimulation (include/fub/RunSimulation.hpp:86-191) restricted to one
 * patch: up to max_steps cycles, stopping at final_time, each step with
 * dt = min(final_time - t, cfl * min_dir stable_dt), the split sweeps in the
 * direction order of MakeSplitDirections (SplittingMethod.hpp:115-126), the
 * per-sweep CFL check of DimensionalSplitLevelIntegrator (:240-245) and
 * restart of a failed step from the untouched step-start state with
 * cfl * failure_dt.  Steps are enqueued `batch` at a time without host
 * synchronisation; the host syncs once per batch.  dt_log (host memory,
 * max_steps doubles, may be NULL) receives the dt of each completed cycle. */
int fvs_level_advance(fvs_level* level, int max_steps, double final_time,
                      int batch, double* dt_log, int* steps_done,
                      int* failures);
/* Time point and cycle count of the current state. */
int fvs_level_time(fvs_level* level, double* t, int64_t* cycles);

/* Speed floor of the step-start CFL pass (no counterpart in the reference,
 * whose ComputeStableDt, HllemMethod.hpp:1152-1216 and
 * GodunovMethod.hpp:163-171, evaluates every face).  The CFL pass of a 3D
 * HLLEM level, and of an exact-Godunov level (2D or 3D, with a Newton
 * iteration cap of at least 16), evaluates the exact speed only of faces
 * whose certified upper bound reaches a floor, the previous pass's largest
 * signal speed less a margin (four times its last relative drop, between
 * 2^-10 and 1/4), and confirms that some evaluated face did reach it.  When
 * none did, the step is repeated with every face evaluated (and the margin
 * quadruples), so the floor never changes a result.
 * fvs_level_hint_max_speed sets the floor for the current state
 * (speed[d] >= 0; 0: none; ignored by other levels); fvs_level_cfl_redos
 * reports how many steps were repeated. */
int fvs_level_hint_max_speed(fvs_level* level, const double speed[3]);
int fvs_level_cfl_redos(fvs_level* level, int64_t* redos);

/* Kernel timing of the level driver (no counterpart in the reference; it
 * replaces reading a profiler).  With timing on, fvs_level_advance brackets
 * every launch on the level's stream with HIP events; fvs_level_kernel_times
 * synchronises the stream, writes the summed milliseconds and launch counts
 * per class since the last read (FVS_LEVEL_TIMERS entries each) and resets
 * them.  Classes: split sweep in x / y / z (the sweep kernel alone) and the
 * step-start CFL pass (reduction and finalize). */
enum fvs_level_timer {
  FVS_TIMER_SWEEP_X = 0,
  FVS_TIMER_SWEEP_Y = 1,
  FVS_TIMER_SWEEP_Z = 2,
  FVS_TIMER_CFL = 3,
  FVS_LEVEL_TIMERS = 4
};
int fvs_level_set_timing(fvs_level* level, int on);
int fvs_level_kernel_times(fvs_level* level, double* ms, int64_t* count);

/* Replay the level's time steps from HIP graphs (on by default; the timing
 * mode launches eagerly): on the first advance with graphs on, every
 * distinct step (per start buffer, cycle parity and attached network) is
 * captured once as a single-step graph and as an 8-step chain, and runs of
 * eight steps are replayed as one graph launch.  No counterpart in the
 * reference, whose steps are host loops; results are identical either way. */
int fvs_level_set_graphs(fvs_level* level, int on);

/* One fused split sweep (flux + update + completion + CFL check of the
 * input) from cons_in to cons_out (distinct arrays, D+2 components of
 * n[0]*n[1]*n[2] each, x fastest, transmissive boundaries).  Writes the
 * minimum face stable dt of the input in dir to *stable_dt_device.  This is
 * the operator one split step of AdvanceLevelNonRecursively applies
 * (DimensionalSplitLevelIntegrator.hpp:227-264). */
int fvs_split_sweep(fvs_handle* h, const double* const* cons_in,
                    double* const* cons_out, const int64_t n[3], double dt,
                    double dx, int dir, double* stable_dt_device);

/* Minimum over directions and faces of the stable dt of a patch with
 * transmissive boundaries (DimensionalSplitLevelIntegrator::ComputeStableDt,
 * :204-219); per-direction minima go to dt_dir_device[0..dim-1]. */
int fvs_patch_stable_dt(fvs_handle* h, const double* const* cons,
                        const int64_t n[3], const double dx[3],
                        double* dt_dir_device);

/* fvs_patch_stable_dt without any allocation: the same kernels and results,
 * with the per-workgroup partial minima in a fixed 2 MiB device array of
 * the library (262,144 keys; a patch needing more returns
 * FVS_BAD_ARGUMENT).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  Every call on one device shares that array, so calls
 * that overlap on one device (different streams or threads) race; calls on
 * one stream are ordered and safe. */
int fvs_patch_stable_dt_noalloc(fvs_handle* h, const double* const* cons,
                                const int64_t n[3], const double dx[3],
                                double* dt_dir_device);

/* HBM streaming probe (no counterpart in the reference; a measurement aid
 * for callers that want the achievable bandwidth of a sweep's traffic):
 * copies ncomp (1..5) device arrays of n doubles (n even, 16-byte aligned),
 * src[c] to dst[c], on the handle's stream; every element is read and written
 * once, 16 bytes per lane and array per access, with nontemporal loads and
 * stores. */
int fvs_stream_copy(fvs_handle* h, const double* const* src, double* const* dst,
                    int ncomp, int64_t n);

/* ---- exact Godunov -------------------------------------------------------- */

/* First-order exact-Godunov split sweep (FluxMethod<Godunov<PerfectGas<D>,
 * ExactRiemannSolver>>, include/fub/flux_method/GodunovMethod.hpp:130-170,
 * src/equations/perfect_gas/ExactRiemannSolver.hpp:34-353) on a ghost-free
 * transmissive patch, fused with its CFL check.  One FP64 Newton iteration
 * per face serves both the flux and the face's signal speeds.
 * guess_device (nullable): per-face initial star pressure in direction dir's
 * face box (n with n[dir] + 1, x fastest); NULL uses 0.5 (pL + pR) as the
 * reference (ExactRiemannSolver.hpp:91).  cons_out may be NULL (CFL only).
 * Newton iterations are added to the handle's counters (fvs_newton_stats). */
int fvs_godunov_sweep(fvs_handle* h, const double* const* cons_in,
                      double* const* cons_out, const int64_t n[3], double dt,
                      double dx, int dir, const double* guess_device,
                      double* stable_dt_device);

/* ExactRiemannSolver::ComputeMiddleState (ExactRiemannSolver.hpp:34-97) on n
 * complete state pairs (device SoA (D+4) x n) along dir, Newton seeded with
 * guess[i] (nullable: 0.5 (pL + pR)).  Outputs (each nullable): star pressure
 * and velocity, Newton iterations, converged flag. */
int fvs_exact_middle_state(fvs_handle* h, int dir, const double* left,
                           const double* right, int64_t n, const double* guess,
                           double* pM, double* uM, int* iterations, int* converged);

/* Copies the handle's Newton counters to the host: stats[0] total
 * iterations, [1] largest per-face count, [2] faces that hit the iteration
 * cap (the reference throws there, NewtonIteration.hpp:44-50), [3] faces,
 * [4] faces whose learned guess did not converge and were redone from the
 * reference's guess.  reset != 0 zeroes them afterwards.  Synchronises. */
int fvs_newton_stats(fvs_handle* h, uint64_t stats[5], int reset);

/* ---- learned Newton initial guess (6 -> 64 -> 64 -> 1 MLP) ----------------
 * No counterpart in the reference: the network predicts
 * r = ln p* - (ln pL + ln pR) / 2 from (ln rhoL, uL/5, ln pL, ln rhoR, uR/5,
 * ln pR); the exact solver's Newton iteration starts from
 * sqrt(pL pR) exp(r) and is otherwise unchanged.  Inference runs on the
 * matrix cores (BF16 operands, FP32 accumulation); training is FP32. */
typedef struct fvs_mlp fvs_mlp;
int fvs_mlp_create(void* hip_stream, uint64_t seed, fvs_mlp** out);
int fvs_mlp_destroy(fvs_mlp* mlp);
int fvs_mlp_num_params(void);
int fvs_mlp_get_params(fvs_mlp* mlp, float* host_params);
int fvs_mlp_set_params(fvs_mlp* mlp, const float* host_params);
/* Generates n labelled pairs on the device: rho, p log-uniform in
 * [1e-3, 1e3], u uniform in [-5, 5]; label = FP64 Newton star pressure.
 * counts[0] usable, [1] vacuum (rejected), [2] not converged (rejected). */
int fvs_mlp_generate(fvs_mlp* mlp, int64_t n, uint64_t seed, double gamma,
                     int64_t counts[3]);
/* Copies the generated set into caller device buffers (each nullable):
 * prim (6 x n: rhoL uL pL rhoR uR pR), features (6 x n), target r (n),
 * flag (n), Newton iterations from the reference guess (n); *n gets the
 * sample count. */
int fvs_mlp_dataset(fvs_mlp* mlp, int64_t* n, double* prim, float* features,
                    float* target, int* flag, int* iterations);
/* FP32 forward of n feature vectors (device, 6 x n) into y (device, n). */
int fvs_mlp_forward(fvs_mlp* mlp, const float* features, int64_t n, float* y);
/* Loss over flag == 0 samples (mean of e^2, or with huber > 0 of the Huber
 * form 2 huber |e| - huber^2 beyond |e| > huber) and its gradient
 * (fvs_mlp_num_params floats, host) on one device batch. */
int fvs_mlp_gradients(fvs_mlp* mlp, const float* features, const float* target,
                      const int* flag, int64_t n, float huber, float* grad_host,
                      float* loss_host);
/* Adam on the generated set, `epochs` passes of minibatches of `batch`,
 * cosine-decayed learning rate from lr, Huber threshold huber (0: MSE). */
int fvs_mlp_train(fvs_mlp* mlp, int epochs, int64_t batch, float lr, float huber,
                  uint64_t seed, float* loss_history);
/* Matrix-core inference on complete state pairs (device SoA (D+4) x n):
 * network output r (nullable) and initial guess (nullable).  Weak pairs
 * (|pL - pR| and |uR - uL| (rhoL + rhoR)(aL + aR) / 4 both below
 * 0.1 (pL + pR)) keep the mean-pressure guess. */
int fvs_mlp_infer_pairs(fvs_mlp* mlp, int dim, int dir, const double* left,
                        const double* right, int64_t n, float* r_out, double* guess);
/* Initial guesses for every face of direction dir of a ghost-free
 * transmissive patch, in the face-box layout fvs_godunov_sweep reads. */
int fvs_mlp_face_guess(fvs_mlp* mlp, int dim, double gamma, const double* const* cons,
                       const int64_t n[3], int dir, double* guess);
/* Attaches a network to a first-order exact-Godunov handle: the level
 * driver then seeds every Newton iteration (fluxes and CFL) with it.
 * NULL detaches.  The handle does not own the network. */
int fvs_set_mlp(fvs_handle* h, fvs_mlp* mlp);

/* ---- KBN h-box cut-cell method (KbnCutCellMethod) ---------------------- */

/* FillCutCellData (src/AMReX/cutcell/FillCutCellData.cpp:28-118): for every
 * face of the box [face_lo, face_hi] normal to dir, the unshielded,
 * shielded-from-left, shielded-from-right and doubly shielded fractions
 * (absolute and relative to beta) from the face fractions of the face and
 * its two neighbours along dir.  Writes geom->unshielded[dir] and friends. */
int fvs_kbn_fill_cutcell_data(fvs_handle* h, const fvs_cutcell_data* geom, int dir,
                              const int64_t face_lo[3], const int64_t face_hi[3]);

/* IsCutCell (src/CutCellData.cpp:225-256) for every cell of [lo, hi]:
 * flags (device int32, x fastest over the box) get 1 for cut cells. */
int fvs_kbn_is_cut_cell(fvs_handle* h, const fvs_cutcell_data* geom, const int64_t lo[3],
                        const int64_t hi[3], int32_t* flags);

/* KbnCutCellMethod::PreAdvanceHierarchy (KbnStabilisation.hpp:188-210):
 * for every cut cell of states' box with positive density, the reference
 * state from the boundary Riemann problem (state reflected at the wall
 * against itself, in the frame rotated onto the boundary normal). */
int fvs_kbn_reference_states(fvs_handle* h, const fvs_view* references,
                             const fvs_view* states, const fvs_cutcell_data* geom);

/* KbnCutCellMethod::ComputeBoundaryFluxes (KbnStabilisation.hpp:212-273),
 * single-block path (reference masks 0): for every cut cell of references'
 * box, F_B = F_dir(reference) + (p_boundary - p_reference) e_dir.  Other
 * cells of boundary_fluxes are left untouched. */
int fvs_kbn_boundary_fluxes(fvs_handle* h, const fvs_view* boundary_fluxes,
                            const fvs_view* states, const fvs_view* references,
                            const fvs_cutcell_data* geom, double dt, double dx, int dir);

/* KbnCutCellMethod::ComputeRegularFluxes (KbnStabilisation.hpp:347-429):
 * the masked flux method over the faces of fluxes' box; faces with beta = 0
 * get a zero flux, slopes next to covered cells are zero. */
int fvs_kbn_regular_fluxes(fvs_handle* h, const fvs_view* fluxes, const fvs_view* states,
                           const fvs_cutcell_data* geom, double dt, double dx, int dir);

/* KbnCutCellMethod::ComputeStableDt (KbnStabilisation.hpp:277-345): minimum
 * of the masked stable dt over the faces of states' box shrunk by the
 * stencil; written to *dt_device. */
int fvs_kbn_stable_dt(fvs_handle* h, const fvs_view* states, const fvs_cutcell_data* geom,
                      double dx, int dir, double* dt_device);

/* KbnCutCellMethod::ComputeCutCellFluxes -> ComputeStableFluxComponents
 * (src/cutcell_method/KbnStabilisation.cpp:55-184): stabilised and singly
 * shielded fluxes over the faces of regular_fluxes' box. */
int fvs_kbn_cutcell_fluxes(fvs_handle* h, const fvs_view* stabilised,
                           const fvs_view* shielded_left, const fvs_view* shielded_right,
                           const fvs_view* boundary_fluxes, const fvs_view* regular_fluxes,
                           const fvs_cutcell_data* geom, double dt, double dx, int dir);

/* cutcell::TimeIntegrator::UpdateConservatively
 * (src/AMReX/cutcell/TimeIntegrator.cpp:65-411) in place on the cells of
 * scratch's box (fluxes on the faces enclosing them), the h-box update with
 * absolute shielded fractions; covered cells become 0. */
int fvs_kbn_update_conservatively(fvs_handle* h, const fvs_view* scratch,
                                  const fvs_view* unshielded, const fvs_view* stabilised,
                                  const fvs_view* shielded_left, const fvs_view* shielded_right,
                                  const fvs_view* boundary_fluxes,
                                  const fvs_cutcell_data* geom, double dt, double dx, int dir);

/* cutcell::TimeIntegrator2::UpdateConservatively
 * (src/AMReX/cutcell/TimeIntegrator.cpp:413-508): dest = src + dt/(dx alpha)
 * (betaL F_L - betaR F_R - (betaL - betaR) F_B) on cells with alpha > 0 and
 * betaL != betaR, the regular update where betaL == betaR; covered cells are
 * left untouched. */
int fvs_kbn_update_conservatively2(fvs_handle* h, const fvs_view* dest, const fvs_view* src,
                                   const fvs_view* stabilised, const fvs_view* boundary_fluxes,
                                   const fvs_cutcell_data* geom, double dt, double dx, int dir);

/* cutcell::Reconstruction::CompleteFromCons, SIMD kernel with volume mask
 * (include/fub/AMReX/cutcell/Reconstruction.hpp:123-157): complete states
 * over complete's box, the zero state where alpha = 0 or rho <= 0. */
int fvs_kbn_complete_from_cons(fvs_handle* h, const fvs_view* complete, const fvs_view* cons,
                               const fvs_cutcell_data* geom);

/* ---- single-patch cut-cell level (2D) --------------------------------------
 * RunSimulation over one 2D patch of n[0] x n[1] cells with two ghost layers
 * and transmissive boundaries, advancing with the reference's cut-cell
 * split sequence: reference states at every step
 * (src/AMReX/cutcell/IntegratorContext.cpp:800-806), then per split step the
 * boundary fill, the KBN CFL check, boundary / regular / stabilised fluxes
 * (include/fub/AMReX/cutcell/FluxMethod.hpp:153-257), the cut-cell update
 * (integrator 0: TimeIntegrator, 1: TimeIntegrator2) and the masked
 * reconstruction.  The handle must be 2D, MUSCL-Hancock with an HLLEM base.
 * geom covers the patch grown by two ghost layers (cells -2 .. n+1, faces
 * -2 .. n+2); the level keeps a dense device copy. */
typedef struct fvs_kbn_level fvs_kbn_level;
int fvs_kbn_level_create(fvs_handle* h, const fvs_patch_desc* patch,
                         const fvs_cutcell_data* geom, int integrator,
                         fvs_kbn_level** out);
int fvs_kbn_level_destroy(fvs_kbn_level* level);
/* Number of interior cut cells. */
int fvs_kbn_level_ncut(fvs_kbn_level* level);
/* Copies 4 interior component arrays (device, n[0]*n[1] each, x fastest). */
int fvs_kbn_level_set_state(fvs_kbn_level* level, const double* const* comps, double t0);
int fvs_kbn_level_get_state(fvs_kbn_level* level, double* const* comps);
int fvs_kbn_level_time(fvs_kbn_level* level, double* t, int64_t* cycle);
/* Same contract as fvs_level_advance. */
int fvs_kbn_level_advance(fvs_kbn_level* level, int max_steps, double final_time,
                          int batch, double* dt_log, int* steps_done, int* failures);

/* ---- batched sweeps over the patches of an AMR level ----------------------
 * All patches of a level (2D, uneven sizes, two ghost layers each) in one
 * pool per state buffer; a sweep is one persistent launch over a tile list
 * of every patch (the reference sweeps patch by patch, ForEachFab,
 * include/fub/AMReX/ForEachFab.hpp:48-57, with FluxMethod::
 * ComputeNumericFluxes and HyperbolicPatchIntegrator::UpdateConservatively
 * per patch).  Handle: 2D MUSCL-Hancock + HLLEM (Larrouturou); precision
 * FVS_FP32_FLUX runs reconstruction and flux in FP32, the update in FP64.
 * Patch p occupies 4 * (nx+4) * (ny+4) doubles from fvs_batch_offset(p):
 * component-major, x fastest, cell (-2, -2) first. */
typedef struct fvs_batch fvs_batch;
/* n: host, 2 per patch (nx, ny); dx: host, one per patch (square cells). */
int fvs_batch_create(fvs_handle* h, int64_t count, const int32_t* n, const double* dx,
                     fvs_batch** out);
int fvs_batch_destroy(fvs_batch* batch);
double* fvs_batch_pool(fvs_batch* batch, int which);
int64_t fvs_batch_pool_size(fvs_batch* batch);
int64_t fvs_batch_offset(fvs_batch* batch, int64_t patch);
int fvs_batch_upload(fvs_batch* batch, int which, const double* src_device);
int fvs_batch_download(fvs_batch* batch, int which, double* dst_device);
/* Time step the sweeps read from device memory. */
int fvs_batch_set_dt(fvs_batch* batch, double dt);
/* Transmissive ghost layers of every patch of pool `which`. */
int fvs_batch_fill_ghosts(fvs_batch* batch, int which);
/* One sweep of direction dir from pool src into pool 1 - src (interior
 * cells); the sweep's CFL check value (nullable) is written to *dt_device:
 * with the time step 0 the minimum HLLEM stable dt of all patch faces; with
 * a positive time step that minimum wherever it lies below the time step
 * (the check fails), and otherwise a value of at least the time step (faces
 * whose certified speed bound proves their stable dt above it are not
 * evaluated exactly). */
int fvs_batch_sweep(fvs_batch* batch, int dir, int src, double* dt_device);
/* Autotuner: times every candidate tile shape of direction dir over `reps`
 * sweeps and keeps the fastest (shape index, per-shape ms; nullable). */
int fvs_batch_tune(fvs_batch* batch, int dir, int reps, int* shape, float* ms);
int fvs_batch_tile_shape(fvs_batch* batch, int dir, int* tx, int* ty, int* ntiles);
/* Selects candidate tile shape `shape` (0..3) of direction dir. */
int fvs_batch_set_shape(fvs_batch* batch, int dir, int shape);
/* Captures one split step (fill 0, X 0 -> 1, fill 1, Y 1 -> 0) into a
 * hipGraph; fvs_batch_replay launches it `steps` times, stream-ordered with
 * the handle's stream.  The sweeps' CFL minima land in fvs_batch_cfl. */
int fvs_batch_capture(fvs_batch* batch);
int fvs_batch_replay(fvs_batch* batch, int steps);
double* fvs_batch_cfl(fvs_batch* batch);
/* Host copy of the two CFL minima (synchronises the handle's stream). */
int fvs_batch_read_cfl(fvs_batch* batch, double* host2);

/* ---- AMR hierarchy over batched levels (2D) -------------------------------
 * Levels 0..nlevels-1 (at most 8) are batches of one handle; level l has dx / ratio^l
 * and a domain of domain[0] x domain[1] cells times ratio^l.  lo: host, the
 * lower corner (i0, j0) of every patch of every level, level 0 first, in
 * the level's cell indices.  Level 0 must cover the domain; patches of
 * finer levels must be aligned to the ratio and properly nested (ghost
 * layers clamped to the domain lie in same-level patches or above
 * coarser-level patches, and the coarse cells beside each patch belong to
 * coarser patches).  Replaces the AMReX machinery the reference drives
 * through IntegratorContext (src/AMReX/IntegratorContext.cpp:380-533):
 * FillGhostLayerTwoLevels / FillGhostLayerSingleLevel (FillPatchTwoLevels
 * with piecewise-constant interpolation; transmissive domain edges),
 * AccumulateCoarseFineFluxes + ApplyFluxCorrection (FluxRegister) and
 * CoarsenConservatively (average_down).  All work goes to the handle's
 * stream; calls on one hierarchy must not overlap (it keeps per-object
 * launch state), and its levels' batches must outlive it. */
typedef struct fvs_amr fvs_amr;
int fvs_amr_create(fvs_handle* h, int nlevels, fvs_batch* const* levels, const int32_t* lo,
                   const int64_t* domain, int ratio, fvs_amr** out);
int fvs_amr_destroy(fvs_amr* amr);
/* Ghost layers of every patch of level `level`, pool `which`: same-level
 * patches first, else the coarser level's cell above (pool 0). */
int fvs_amr_fill_ghosts(fvs_amr* amr, int level, int which);
/* Cells of level fine_level - 1 (pool 0) under level fine_level's patches
 * become the mean of their ratio^2 fine cells (pool 0). */
int fvs_amr_average_down(fvs_amr* amr, int fine_level);
/* One coarse time step dt of the hierarchy (SubcycleFineFirstSolver::
 * AdvanceHierarchy, include/fub/solver/SubcycleFineFirstSolver.hpp:114-175):
 * level l+1 advances ratio times with dt / ratio before level l; each level
 * step is the two split sweeps (order per MakeSplitDirections,
 * include/fub/split_method/SplittingMethod.hpp:115-126) with ghost layers
 * refilled before each sweep; then reflux and average-down onto level l.
 * States live in pool 0 of each level between steps; pool 1 is scratch. */
int fvs_amr_advance(fvs_amr* amr, double dt);
/* Per level, the CFL check value of the sweeps of the last advance (host,
 * nlevels values; synchronises): the minimum stable dt over the level's
 * sweeps wherever it lies below the level's own time step dt / ratio^l
 * (the reference's TimeStepTooLarge, whose retry uses that minimum), and
 * otherwise a value of at least that time step (see fvs_batch_sweep). */
int fvs_amr_level_cfl(fvs_amr* amr, double* host);
/* Stable coarse dt of the current state: min over levels of ratio^l times
 * the level's minimum HLLEM stable dt over both directions
 * (SubcycleFineFirstSolver::ComputeStableDt, :101-112), without the CFL
 * factor.  Uses pool 1 of each level as scratch; synchronises. */
int fvs_amr_stable_dt(fvs_amr* amr, double* dt_host);
/* Coarse-step count of the hierarchy (it selects the split order of the
 * next step, MakeSplitDirections).  A caller that undoes a step by
 * restoring the levels' states restores the count with it, as the
 * reference restores its whole solver from the backup
 * (include/fub/RunSimulation.hpp:141-151). */
int fvs_amr_cycle(fvs_amr* amr, int64_t* cycle);
int fvs_amr_set_cycle(fvs_amr* amr, int64_t cycle);
/* Capture each coarse step (per cycle parity) once into a HIP graph and
 * replay it (on by default; results are identical either way).  The step's
 * dt is read from device memory, so one graph serves every dt; retuning a
 * level's tiles triggers a new capture.  No counterpart in the reference. */
int fvs_amr_set_graphs(fvs_amr* amr, int on);
/* Ghost-layer sequencing of a level step.  FVS_AMR_SEQ_REFERENCE (default):
 * the ghost layers are filled once at the start of the level step
 * (DimensionalSplitLevelIntegrator::PreAdvanceLevel,
 * include/fub/solver/DimensionalSplitLevelIntegrator.hpp:169-190), the first
 * split sweep also advances the two ghost rows across its direction (the
 * grown scratch, include/fub/AMReX/IntegratorContext.hpp:52-53), and before
 * the second direction only that direction's physical boundary is applied
 * (ApplyBoundaryCondition, :230-233, TransmissiveBoundary).
 * FVS_AMR_SEQ_REFILL: every split sweep refills all ghost layers. */
#define FVS_AMR_SEQ_REFERENCE 0
#define FVS_AMR_SEQ_REFILL 1
int fvs_amr_set_sequencing(fvs_amr* amr, int mode);
/* Refinement tags of a level's current state (pool 0, ghosts refilled
 * first): ScalarGradientDetector::TagCellsForRefinement
 * (include/fub/tagging_method/GradientDetector.hpp:150-194) with the
 * density and pressure projections and their tolerances.  tags: device,
 * one byte per interior cell (1 tagged), patches in level order, x fastest
 * within a patch (sum of nx * ny bytes). */
int fvs_amr_tag_cells(fvs_amr* amr, int level, double tol_rho, double tol_p, uint8_t* tags);
/* State of hierarchy src moved onto hierarchy dst (same handle, domain,
 * ratio and number of levels; typically dst has regridded finer levels),
 * level by level into pool 0: a cell takes src's same-level cell where src
 * has one, otherwise dst's coarser-level cell above it (piecewise
 * constant), the fill after GriddingAlgorithm::RegridAllFinerlevels
 * (src/AMReX/GriddingAlgorithm.cpp:86-92, FillPatchTwoLevels with
 * pc_interp).  Synchronises; allocates its link lists temporarily. */
int fvs_amr_transfer(fvs_amr* dst, fvs_amr* src);

#ifdef __cplusplus
}
#endif

#endif /* FVS_H */
