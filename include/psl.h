/*
 * psl.h -- C ABI of the MI355X-native PeriodicShiftLBM D3Q27 step.
 *
 * Plain C: opaque handles, plain pointers and sizes.  Every function returns
 * 0 on success or a negative PSL_E* code; psl_last_error() returns a
 * thread-local message for the last failure.  Device memory is owned by the
 * library; host arrays passed in are only read (or written) during the call.
 * Calls on one grid are ordered on that grid's HIP stream; only functions that
 * return data to the host synchronise it.  One grid must not be driven from
 * two threads at once; distinct grids are independent.
 *
 * Each entry point names the reference interface it replaces
 * (stepanmuller/PeriodicShiftLBM, paths relative to the repository root).
 */
#ifndef PSL_H
#define PSL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSL_ABI_VERSION 1

enum {
  PSL_OK = 0,
  PSL_E_ARG = -1,       /* bad argument / size */
  PSL_E_HIP = -2,       /* HIP runtime error */
  PSL_E_STATE = -3,     /* call not valid in the current state */
  PSL_E_IO = -4,        /* file could not be read */
  PSL_E_NOMEM = -5,     /* allocation failed */
  PSL_E_GEOMETRY = -6   /* mesh / interface check failed */
};

/* storage of the distributions */
enum {
  PSL_F32 = 0,      /* f_q in FP32 (216 B per cell update) */
  PSL_F16DEV = 1,   /* f_q - w_q in FP16, collided in FP32 (108 B) */
  PSL_BF16DEV = 2   /* f_q - w_q in BF16, collided in FP32 (108 B) */
};

/* cell flag byte: kind in bits 0-2, periodic x/y/z in bits 3-5 (that outer
 * normal component is ignored), given-value slot in bits 6-7.
 * Replaces MarkerStruct (include/types.h:79-80) and the getMarkers callback. */
enum {
  PSL_FLUID = 0,
  PSL_BOUNCEBACK = 1,
  PSL_GIVEN_RHO = 2,   /* givenRho: restoreUxUyUz + applyMBBC */
  PSL_GIVEN_U = 3,     /* givenUxUyUz: restoreRho + applyMBBC */
  PSL_GIVEN_NONE = 4,  /* neither: restoreRhoUxUyUz + applyMBBC */
  PSL_GIVEN_RHOU = 5,  /* both: applyMBBC */
  PSL_GHOST = 6,       /* Marker.ghost: not updated */
  PSL_COVERED = 7,     /* inside the finer grid's sub-box: not updated */
  PSL_PERIODIC_X = 8,
  PSL_PERIODIC_Y = 16,
  PSL_PERIODIC_Z = 32
};
#define PSL_FLAG(kind, slot) ((uint8_t)((kind) | ((slot) << 6)))

/* Mirrors InfoStruct (include/types.h:73-75). */
typedef struct psl_info {
  float gridID;
  float res;
  int cellCountX, cellCountY, cellCountZ, cellCount;
  float ox, oy, oz;
  float nu;
  float dtPhys;
  int iSubgridStart, jSubgridStart, kSubgridStart;
  int iSubgridEnd, jSubgridEnd, kSubgridEnd;
  int cellCountFineToCoarse, cellCountCoarseToFine;
  float pRegulator, iRegulator;
} psl_info;

typedef struct psl_grid psl_grid;
typedef struct psl_stl psl_stl;
typedef struct psl_mask psl_mask;

const char* psl_last_error(void);
int psl_abi_version(void);
/* HIP device count visible to the library (0 when none). */
int psl_device_count(int* count);

/* ---------------------------------------------------------------- grids */
/* GridStruct allocation (include/types.h:77; flowAroundSphere/flowAroundSphereMainSingle.cu:180-181):
 * fArray(27, cellCount) and shifter = 0. cellCount is recomputed from the counts. */
int psl_grid_create(const psl_info* info, int storage, psl_grid** out);
int psl_grid_destroy(psl_grid* grid);
int psl_grid_get_info(const psl_grid* grid, psl_info* out);
/* Sub-grid bounds and nu may change after creation (the driver writes Info fields
 * between levels, flowAroundSphereMainSingle.cu:199-204). */
int psl_grid_set_info(psl_grid* grid, const psl_info* info);
int psl_grid_set_stream(psl_grid* grid, void* hip_stream);
/* Per-cell flags (N bytes, host), or NULL for all fluid: the host-evaluated
 * getMarkers / bouncebackMarkerArray (include/applyLocalCellUpdate.h:5-23). */
int psl_grid_set_flags(psl_grid* grid, const uint8_t* host_flags);
int psl_grid_get_flags(const psl_grid* grid, uint8_t* host_flags);
/* getGivenRhoUxUyUz table: 4 slots x (rho, ux, uy, uz) (flowAroundSphereMainSingle.cu:69-77). */
int psl_grid_set_bc_values(psl_grid* grid, const float* values16);
/* getSmagorinskyConstant (flowAroundSphereMainSingle.cu:79-82): a constant, or a
 * per-cell host array of N floats when per_cell != NULL. */
int psl_grid_set_smagorinsky(psl_grid* grid, float cs, const float* per_cell);
/* Bit 0: apply applyBCPeriodicX, bit 1: applyBCPeriodicY inside psl_update_grid. */
int psl_grid_set_periodic_copies(psl_grid* grid, int mask);

/* fillEquilibriumFromFunction (include/fillEquilibrium.h:1-22): f = feq(rho, u) from
 * host arrays of N values; NULL arrays mean rho = 1 / u = 0. */
int psl_fill_equilibrium(psl_grid* grid, const float* rho, const float* ux, const float* uy, const float* uz);

/* applyStreaming(GridStruct&) (include/applyStreaming.h:11-31): shifter update only. */
int psl_apply_streaming(psl_grid* grid);
/* applyLocalCellUpdate(GridStruct&) (include/applyLocalCellUpdate.h:1-74). */
int psl_apply_local_cell_update(psl_grid* grid);
/* applyBCPeriodicX / applyBCPeriodicY (include/boundaryConditions/periodicXY.h:38-128). */
int psl_apply_bc_periodic_x(psl_grid* grid);
int psl_apply_bc_periodic_y(psl_grid* grid);

/* applyCoarseFineGridCommunication (include/gridRefinementFunctions.h:433-526). */
int psl_coarse_fine_communication(psl_grid* coarse, psl_grid* fine);
/* fillCoarseGridFromFine (include/gridRefinementFunctions.h:528-536). */
int psl_fill_coarse_from_fine(psl_grid* coarse, psl_grid* fine);
/* updateGrid recursion (flowAroundSphere/flowAroundSphereMainSingle.cu:158-167),
 * nsteps coarse steps over levels[0..nlevels-1]. */
int psl_update_grid(psl_grid* const* levels, int nlevels, int nsteps);

/* Shifter (include/types.h:77 GridStruct.shifter). */
int psl_get_shifter(const psl_grid* grid, int* out27);
int psl_set_shifter(psl_grid* grid, const int* in27);
/* Distributions in logical order: host[q*N + cell] = f_q(cell) as FP32. */
int psl_get_f(psl_grid* grid, float* host);
int psl_set_f(psl_grid* grid, const float* host);
/* rho, ux, uy, uz per cell (getRhoUxUyUz, include/cellFunctions.h:135-155); any may be NULL. */
int psl_get_macroscopic(psl_grid* grid, float* rho, float* ux, float* uy, float* uz);
/* The same fields into a device buffer of 4 x cellCount floats (rho, ux, uy,
 * uz), ordered on the grid's stream, no host copy. */
int psl_get_macroscopic_device(psl_grid* grid, float* out_device);
/* Device pointer of fArray (storage element type), for callers that keep data resident. */
int psl_grid_device_f(psl_grid* grid, void** dev_ptr, size_t* bytes);

/* Momentum exchange over bounceback cells of the box [lo, hi): out = sum rho (u - u_bounced),
 * accumulated in FP64 (getSphereDrag, flowAroundSphereMainSingle.cu:99-156; getLift,
 * NACA0012/NACA0012Main.cu:96-133).  lo/hi may be NULL for the whole grid. */
int psl_reduce_momentum_exchange(psl_grid* grid, const int* lo3, const int* hi3, double* out3);
/* psl_reduce_momentum_exchange without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * per-step drag loops, harnesses that forbid hipMalloc).  Its per-block partials live in a
 * fixed per-device array of PSL_NOALLOC_MAX_BOX_CELLS / 256 * 3 + 4 doubles (1.5 MiB),
 * held until the process ends; a box of more than PSL_NOALLOC_MAX_BOX_CELLS cells is
 * rejected with PSL_E_ARG.  Same kernels and block count as psl_reduce_momentum_exchange,
 * so the same bits.  Call with the grid's device current.  Every call on one device
 * shares that device's array: calls that overlap on one device (other host threads,
 * other streams) race. */
#define PSL_NOALLOC_MAX_BOX_CELLS 16777216
#define PSL_NOALLOC_MAX_DEVICES 64
int psl_reduce_momentum_exchange_noalloc(psl_grid* grid, const int* lo3, const int* hi3, double* out3);
/* Total mass sum f over updated cells (not ghost / covered / bounceback), FP64. */
int psl_reduce_mass(psl_grid* grid, double* out);

/* Wait for all work queued on the device (hipDeviceSynchronize). */
int psl_device_synchronize(void);
/* Bulk-kernel timing: while enabled, every fused bulk update of this grid is
 * bracketed by HIP events on the grid's stream; read returns the summed kernel
 * time in ms and the launch count since enabling (this synchronises). */
int psl_grid_profile_enable(psl_grid* grid, int enable);
int psl_grid_profile_read(psl_grid* grid, double* total_ms, int* launches);
/* Plain device copy of `bytes` bytes (16-byte vector loads/stores): a kernel of
 * known traffic, used to calibrate hardware byte counters against the step kernels. */
/* The HBM rate (read + write GB/s) of the periodic-shift access pattern alone
 * on an nx x ny x nz grid: every cell reads its 27 shifted slots and writes
 * them back, no collision, `launches` times.  The data-movement ceiling of the
 * local update at that size and storage (no reference counterpart). */
int psl_probe_ps_pattern(int nx, int ny, int nz, int storage, int launches, double* gbs);
/* The same at a cap of `workgroups_per_cu` workgroups per CU (0: none; 2-6 as
 * psl_grid_set_occupancy takes), the ceiling of a bulk kernel that runs under
 * that cap (no reference counterpart). */
int psl_probe_ps_pattern_capped(int nx, int ny, int nz, int storage, int workgroups_per_cu, int launches,
                                double* gbs);
/* The same access pattern taken apart (no collision, on a scratch allocation of
 * 27 rows of nx*ny*nz elements): `streams` populations (1, 2, 4, 8, 16 or 27); mode 0 reads then
 * writes in place, 1 only reads, 2 only writes; cells_per_lane 1, 2 or 4 consecutive
 * cells per lane as one access (16-bit storage pairs cells only with
 * line_aligned); line_aligned rounds the shifts down to whole 256-byte runs.
 * gbs = bytes moved per launch / mean launch time. */
int psl_probe_ps_access(int nx, int ny, int nz, int storage, int streams, int mode, int cells_per_lane,
                        int line_aligned, int launches, double* gbs);
int psl_device_copy(void* dst, const void* src, size_t bytes);
/* Device buffer helpers for callers without a device allocator. */
int psl_device_alloc(size_t bytes, void** out);
int psl_device_free(void* ptr);
/* Synchronous host <-> device copies for such buffers. */
int psl_device_upload(void* dst_device, const void* src_host, size_t bytes);
int psl_device_download(void* dst_host, const void* src_device, size_t bytes);
/* Local-update variants (bits): non-temporal loads / stores and XCD-contiguous
 * block order in the bulk kernel; FUSED_BOUNDARY updates the box-face list in
 * the same launch as the bulk instead of a second launch. */
enum {
  PSL_TUNE_NT_LOAD = 1,
  PSL_TUNE_NT_STORE = 2,
  PSL_TUNE_XCD_REMAP = 4,
  PSL_TUNE_FUSED_BOUNDARY = 8,
  PSL_TUNE_VARIANTS = 16
};
int psl_grid_set_tuning(psl_grid* grid, int tuning);
int psl_grid_get_tuning(const psl_grid* grid, int* tuning);
/* Cap on the workgroups of the bulk update kernel that share a compute unit:
 * 0 (no cap: the kernel's register budget allows 7) or 2-6.  The launch then
 * reserves that share of the CU's LDS (the kernel uses none), so fewer
 * workgroups run per CU at once and the HBM sees fewer row streams in flight.
 * Results are the same bits at every cap; only the speed changes.  No
 * reference counterpart (a launch-shape setting, like the variants above). */
int psl_grid_set_occupancy(psl_grid* grid, int workgroups_per_cu);
int psl_grid_get_occupancy(const psl_grid* grid, int* workgroups_per_cu);
/* Autotuner: times all PSL_TUNE_VARIANTS variants on this grid (`steps` whole
 * steps each after one warm-up: streaming, the grid's periodic copies and the
 * local update, so the rows are misaligned as in a run; the state and the
 * shifter advance); the variants within 2 % of the fastest are timed again
 * over max(steps, 256) steps (the time per step runs in phases tens of steps
 * long as the shifts carry the rows across lines and channels) and the
 * fastest of those is kept;
 * then, for a variant whose bulk kernel is a launch of its own (every variant
 * on a grid without listed boundary cells), times the
 * occupancy caps 5, 4 and 3 over as many steps and keeps one if it is faster
 * than no cap by more than 0.5 % (psl_grid_get_occupancy tells which); best and
 * times_ms[PSL_TUNE_VARIANTS] (mean ms per step, no cap) are optional. */
int psl_grid_autotune(psl_grid* grid, int steps, int* best, double* times_ms);
/* Mean duration of the grid's dominant update kernel (kernel = 0: k_ps_bulk,
 * 1: k_ps_fused, 2: the bulk kernel with the learned closure), launched
 * `launches` times back to back between two events on the grid's stream.  The
 * kernel updates the distributions each time (no streaming in between), so
 * call it where the state no longer matters. */
int psl_grid_time_update_kernel(psl_grid* grid, int launches, double* ms_per_launch, int* kernel);
/* The same with psl_apply_streaming before every launch (a host-side shift
 * update, no kernel): the launches see the shifts of consecutive steps, whose
 * alignment to cache lines and memory channels changes the duration by up to
 * ~10 % from step to step, so the mean is that of a run's update kernels.  The
 * shifter advances by `launches` (no reference counterpart). */
int psl_grid_time_update_kernel_streaming(psl_grid* grid, int launches, double* ms_per_launch, int* kernel);
/* Cells handled by the bulk kernel (fluid + bounceback) and by the boundary list. */
int psl_grid_cell_census(const psl_grid* grid, int64_t* bulk_cells, int64_t* boundary_cells, int64_t* skipped_cells);

/* ------------------------------------------------ restart state, health */
/* No reference counterpart (the reference never writes fArray and has no guard against
 * a diverging run).  The restart state is a 256-byte header (magic "PSLSTATE", version,
 * kind, storage, sizes, shifter, closure step count) followed by the storage words of f
 * exactly as the device holds them (FP32, or the 16-bit f - w words), then the scalar
 * populations (27 x N FP32) if psl_scalar_enable was called and the held network output
 * (N FP16) if the grid evaluates a learned closure every k > 1 steps.  A grid with the same
 * info, storage and settings loaded from it continues bit for bit.  Flags, given values
 * and Cs are configuration, not state.  bytes must equal psl_grid_state_bytes; loading
 * checks the header against the grid (PSL_E_ARG on any mismatch). */
int psl_grid_state_bytes(const psl_grid* grid, int64_t* bytes);
int psl_grid_save_state(psl_grid* grid, void* host, int64_t bytes);
int psl_grid_load_state(psl_grid* grid, const void* host, int64_t bytes);
/* Cells with a non-finite population or a density <= 0 (density = the sum of the cell's
 * populations as the next local update reads them), and the least and greatest density
 * of the other cells (+inf / -inf when there are none); rho_min / rho_max may be NULL. */
int psl_grid_check_health(psl_grid* grid, int64_t* bad_cells, double* rho_min, double* rho_max);

/* ------------------------------------------------------- batched grids */
/* Many independent grids advanced together (no reference counterpart: the
 * reference steps one grid at a time).  Each grid keeps its own flags, given
 * values, Cs (uniform or per cell) and storage (all grids share the storage);
 * one step of all grids is a device shift update, one bulk launch over the
 * blocks of every grid and one launch over all boundary lists (or one fused
 * launch).  Steps are captured in hipGraphs of one step and of
 * steps_per_graph steps (default 1000) and replayed, so nsteps steps take
 * nsteps / steps_per_graph + nsteps % steps_per_graph graph launches.  Every
 * grid's result is bit-identical to stepping it alone.  Grids need flags and
 * must not use periodic copies, sub-grids or a learned closure (checked at
 * every psl_batch_step).  The grids' own handles stay valid (get_f,
 * reductions, single-grid steps); changes to a grid's flags, values, nu or Cs
 * are picked up at the next psl_batch_step.  Destroy the batch before its
 * grids. */
typedef struct psl_batch psl_batch;
int psl_batch_create(psl_grid* const* grids, int count, psl_batch** out);
int psl_batch_destroy(psl_batch* batch);
int psl_batch_step(psl_batch* batch, int nsteps);
/* Steps captured in the multi-step graph, 1 .. PSL_BATCH_MAX_STEPS_PER_GRAPH (1: one
 * graph launch per step). */
#define PSL_BATCH_MAX_STEPS_PER_GRAPH 10000
int psl_batch_set_steps_per_graph(psl_batch* batch, int steps);
int psl_batch_get_steps_per_graph(const psl_batch* batch, int* steps);
/* order 0: grids in the given order, 1: largest grid first (block order of the launch). */
int psl_batch_set_tuning(psl_batch* batch, int tuning, int order);
int psl_batch_get_tuning(const psl_batch* batch, int* tuning, int* order);
/* Times the PSL_TUNE_VARIANTS variants x 2 orders (one warm-up step and `steps`
 * steps each; the grids advance) and keeps the fastest; times_ms[2 * PSL_TUNE_VARIANTS]
 * receives ms per step at order * PSL_TUNE_VARIANTS + variant. */
int psl_batch_autotune(psl_batch* batch, int steps, int* best_tuning, int* best_order, double* times_ms);
/* Layout of the two-launch bulk kernel: cells_per_thread 1 or 2 (a workgroup updates
 * 256 or 512 consecutive cells of one grid), tile_order 0 (each grid's tiles in turn)
 * or 1 (interleaved across grids: tile t of every grid that has one, then t + 1).  Cells
 * are stored in the reference's order (k, j, i) whatever the layout: periodic-shift
 * streaming shifts the linear cell index, so the tiles, not the cells, are reordered.
 * Results are bit-identical at every layout. */
int psl_batch_set_layout(psl_batch* batch, int cells_per_thread, int tile_order);
int psl_batch_get_layout(const psl_batch* batch, int* cells_per_thread, int* tile_order);
/* psl_batch_autotune at the four layouts (the fused variants, which keep their own
 * workgroup order, at layout 1/0 only) and keeps the fastest configuration:
 * best[4] = {variant, order, cells_per_thread, tile_order}; times_ms (optional,
 * PSL_BATCH_LAYOUT_CONFIGS entries) receives ms per step at
 * ((tile_order * 2 + cells_per_thread - 1) * 2 + order) * PSL_TUNE_VARIANTS + variant,
 * -1 where a configuration is not run. */
#define PSL_BATCH_LAYOUT_CONFIGS (8 * PSL_TUNE_VARIANTS)
int psl_batch_autotune_layouts(psl_batch* batch, int steps, int best[4], double* times_ms);
int psl_batch_census(const psl_batch* batch, int64_t* cells, int* blocks, int* boundary_cells);

/* ----------------------------------------------- learned LES closure */
/* No reference counterpart: replaces the Smagorinsky rate of getOmegaLES
 * (include/cellFunctions.h:157-198) in the bulk cells by
 *   tau_eff = 3 (nu + 1e-3 * max(0, y)) + 1/2,
 * y = network(standardised features of the non-equilibrium second moments);
 * network 9 -> 32 tanh -> 32 tanh -> 1, evaluated with bf16 matrix cores in
 * the step.  Box-face boundary-list cells keep Smagorinsky (the fallback).
 * Flat parameter vector (PSL_MLP_PARAMS floats): W1[9][32], b1[32],
 * W2[32][32], b2[32], W3[32], b3.  Raw feature rows are 9 floats per cell:
 * |dev Pi|/rho, tr Pi/rho, dev Pi/|dev Pi| (xx, yy, zz, xy, xz, yz),
 * 3 sqrt(6) det(dev Pi)/|dev Pi|^3, with Pi = kappa - rho/3 I. */
enum { PSL_MLP_INPUTS = 9, PSL_MLP_PARAMS = 1409 };
typedef struct psl_mlp psl_mlp;
int psl_mlp_create(psl_mlp** out);
int psl_mlp_destroy(psl_mlp* mlp);
/* mean / std: input standardisation (NULL keeps the current one). */
int psl_mlp_set_params(psl_mlp* mlp, const float* params, const float* mean, const float* std);
int psl_mlp_get_params(const psl_mlp* mlp, float* params, float* mean, float* std);
/* FP32 training kernels on standardised inputs x_dev (n x 9, device) and
 * targets t_dev (device): mean squared error and its gradient (host arrays). */
int psl_mlp_loss_grad(const psl_mlp* mlp, const float* x_dev, const float* t_dev, int n, double* loss,
                      float* grad);
/* Adam training (iters steps, mini-batches of `batch` rows drawn from seed) on
 * raw feature rows raw_dev (n x 9) and targets t_dev (nu_t / 1e-3); the
 * standardisation is taken from the data.  loss_hist[iters] is optional. */
int psl_mlp_train(psl_mlp* mlp, const float* raw_dev, const float* t_dev, int n, int batch, int iters, float lr,
                  uint64_t seed, float* loss_hist);
/* Network output y of raw feature rows (device arrays) with the step's bf16 path. */
int psl_mlp_eval(const psl_mlp* mlp, const float* raw_dev, int n, float* y_dev);
/* Training data from a fine periodic run.  psl_les_filter: box filter of width
 * r over the unique n^3 region of the fine macroscopic fields (device arrays
 * in cell order of a fine_nx x fine_ny x * grid) -> coarse[10 * m^3], m = n/r:
 * rho, Favre u (3), subgrid stress (xx, yy, zz, xy, xz, yz).
 * psl_les_targets: coarse strain by periodic central differences, target
 * nu_t* / 1e-3 = max(0, -dev(tau):S / (2 S:S)) / 1e-3, and the raw features
 * of Pi = -(2/3) rho (3 (nu + nu_t*) + 1/2) S (feat: m^3 x 9). */
int psl_les_filter(const float* rho, const float* ux, const float* uy, const float* uz, int fine_nx, int fine_ny,
                   int n, int r, float* coarse);
int psl_les_targets(const float* coarse, int m, float nu, float* feat, float* target);
/* Use the network in the grid's bulk update (NULL: back to Smagorinsky).  The
 * network must outlive its use; the fused-boundary variant is not used. */
int psl_grid_set_closure(psl_grid* grid, const psl_mlp* mlp);
/* Evaluate the grid's network every `steps` local updates (default 1).  The evaluating
 * update stores each cell's network output as FP16 and relaxes with that rounded value;
 * the updates between read it back and relax with the same eddy viscosity (the network's
 * cost is paid once per `steps`).  The next update after this call evaluates. */
int psl_grid_set_closure_every(psl_grid* grid, int steps);
/* The held network output per cell (FP32 host array of cellCount values). */
int psl_grid_get_closure_y(psl_grid* grid, float* host);

/* --------------------------------------------- esotwist (DIAD) grids */
/* DIADGridStruct (include/types.h:166-171) from host cell tables: ijk3 = 3 arrays of n
 * (i, j, k), nbr7 = 7 arrays of n neighbour cell indices (+i, +j, +k, +ij, +ik, +jk, +ijk). */
typedef struct psl_diad psl_diad;
int psl_diad_create(const psl_info* info, int storage, int64_t n, const int* ijk3, const int* nbr7, psl_diad** out);
int psl_diad_destroy(psl_diad* grid);
int psl_diad_set_flags(psl_diad* grid, const uint8_t* host_flags);
int psl_diad_set_bc_values(psl_diad* grid, const float* values16);
int psl_diad_set_smagorinsky(psl_diad* grid, float cs);
int psl_diad_set_stream(psl_diad* grid, void* hip_stream);
/* fillEquilibriumFromFunction(DIADGridStruct&) (include/fillEquilibrium.h:25-70). */
int psl_diad_fill_equilibrium(psl_diad* grid, const float* rho, const float* ux, const float* uy, const float* uz);
/* applyStreaming(DIADGridStruct&) (include/applyStreaming.h:44-47): flips the parity. */
int psl_diad_apply_streaming(psl_diad* grid);
/* applyLocalCellUpdate(DIADGridStruct&) (include/applyLocalCellUpdate.h:77-171). */
int psl_diad_apply_local_cell_update(psl_diad* grid);
int psl_diad_get_flipper(const psl_diad* grid, int* out);
/* Post-collision populations per cell through the write maps of the current parity
 * (getEsotwistWriteIndex, include/applyStreaming.h:138-192): host[q*n + cell]. */
int psl_diad_get_f(psl_diad* grid, float* host);
int psl_diad_set_f(psl_diad* grid, const float* host);
/* Cell tables of a dense nx*ny*nz box; periodic_mask bits 0/1/2 wrap x/y/z. */
int psl_diad_dense_box(int nx, int ny, int nz, int periodic_mask, int* ijk3, int* nbr7);

/* ------------------------------------------------ DIAD grid builder */
/* buildDIADGrids (include/DIADFunctions.h:640-1046): automatic wall-refined
 * levels on the device.  Level 0 starts as the whole base box; each coarser
 * level keeps its fluid cells, their border and the coarse<->fine interface,
 * and hands a region around every finest-resolution wall cell (widened by
 * wall_refinement_span + levels - level - 1 layers and edited by the
 * refinement rules) to the next level as 2x2x2 children.
 *
 * The drivers' getMarkers callbacks (evaluated by the reference on the device
 * during the build) are given as an ordered rule list: a rule applies to the
 * levels in level_mask (bit = gridID) and, where its shape contains the cell,
 * sets the queried marker (`what`) to `value`; later rules win.  Positions are
 * iCell * res + ox etc. in FP32.  Shapes (p: floats, idx: ints):
 *   SPHERE_LE / SPHERE_LT: |x - p[0..2]|^2 <= / < p[3]^2
 *   XCYLINDER: 0 < x - p[0] < p[4] and (y - p[1])^2 + (z - p[2])^2 < p[3]^2
 *   BOX: p[0] < x < p[3], p[1] < y < p[4], p[2] < z < p[5]
 *   INDEX_BOX: idx[0] <= i <= idx[1], idx[2] <= j <= idx[3], idx[4] <= k <= idx[5];
 *              a negative bound counts from the level's box end (-1 = last). */
enum { PSL_RULE_BOUNCEBACK = 0, PSL_RULE_REFINEMENT = 1, PSL_RULE_OUTLET = 2 };
enum { PSL_SHAPE_SPHERE_LE = 0, PSL_SHAPE_SPHERE_LT = 1, PSL_SHAPE_XCYLINDER = 2, PSL_SHAPE_BOX = 3,
       PSL_SHAPE_INDEX_BOX = 4 };
typedef struct psl_marker_rule {
  int what, value;
  unsigned level_mask;
  int shape;
  float p[8];
  int idx[6];
} psl_marker_rule;
/* Per-level arrays (counts in elements): */
enum {
  PSL_DIAD_IJK = 0,        /* int, 3 x n: i, j, k, cells sorted by (k, j, i) */
  PSL_DIAD_BOUNCEBACK = 1, /* uint8, n: bouncebackMarkerArray */
  PSL_DIAD_INTERFACE = 2,  /* uint8, n: interfaceMarkerArray (coarse levels) */
  PSL_DIAD_NBR = 3,        /* int, 10 x n: EsotwistNbrArray +i +j +k +ij +ik +jk +ijk -i -j -k */
  PSL_DIAD_C2F_READ = 4,   /* int: coarseToFineReadArray (coarse cells) */
  PSL_DIAD_C2F_WRITE = 5,  /* int: coarseToFineWriteArray (fine cells) */
  PSL_DIAD_F2C_WRITE = 6,  /* int: fineToCoarseWriteArray (coarse cells) */
  PSL_DIAD_F2C_READ = 7,   /* int: fineToCoarseReadArray (fine cells) */
  PSL_DIAD_OUTLET = 8,     /* int: outletCellArray */
  PSL_DIAD_ARRAYS = 9
};
typedef struct psl_diad_levels psl_diad_levels;
/* base: level-0 Info (res, ox, oy, oz, nu, dtPhys, cellCountX/Y/Z); finer levels
 * get res/2, dtPhys/2, origin - res_fine/2, doubled counts and
 * nu = dtPhys * nu_phys / (res / 1000)^2.  stls may be NULL when nstl = 0. */
int psl_diad_build(const psl_info* base, int levels, int wall_refinement_span, float nu_phys,
                   const psl_stl* const* stls, int nstl, const psl_marker_rule* rules, int nrules,
                   psl_diad_levels** out);
int psl_diad_levels_destroy(psl_diad_levels* levels);
/* Level Info (cellCount = kept cells) and the element count of every array. */
int psl_diad_levels_info(const psl_diad_levels* levels, int level, psl_info* info, int64_t counts[PSL_DIAD_ARRAYS]);
int psl_diad_levels_get(const psl_diad_levels* levels, int level, int which, void* host);
/* Build time of the last psl_diad_build (ms, device work and host sync included). */
int psl_diad_levels_build_ms(const psl_diad_levels* levels, double* ms);
/* An esotwist grid of one built level, made on the device: cell tables, the
 * first seven neighbour columns, flags BOUNCEBACK / FLUID from the level's
 * bounceback marker (psl_diad_set_flags overrides them) and, below the finest
 * level, the interface tables.  Equivalent to the reference's DIADGridStruct
 * after buildDIADGrids. */
int psl_diad_create_level(const psl_diad_levels* levels, int level, int storage, psl_diad** out);

/* ---------------------------------------- DIAD levels: transfers and steps */
/* coarseToFine{Read,Write}Array / fineToCoarse{Write,Read}Array of a coarse
 * DIADGridStruct (include/types.h:168-169) from host arrays. */
int psl_diad_set_interface(psl_diad* coarse, const psl_diad* fine, const int* c2f_read, const int* c2f_write,
                           int64_t n_c2f, const int* f2c_write, const int* f2c_read, int64_t n_f2c);
/* applyCoarseFineGridCommunication(DIADGridStruct&, DIADGridStruct&)
 * (include/gridRefinementFunctions.h:540-708): fine -> coarse means, then
 * coarse -> fine injection, both rescaled (rescaleF).  The two grids must share
 * the storage type and the stream. */
int psl_diad_coarse_fine_communication(psl_diad* coarse, psl_diad* fine);
/* updateGrid over DIAD levels (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:200-209):
 * streaming + local update of level l, two steps of level l+1, then the
 * communication of l and l+1; nsteps coarse steps. */
int psl_diad_update_grid(psl_diad* const* levels, int nlevels, int nsteps);
/* Interface records of a level: psl_diad_update_grid keeps, for the cells the level
 * transfers read, a contiguous copy of the populations a local update leaves them with,
 * and the transfers read those instead of 27 scattered slots per cell.  Child records
 * (the children of the coarser level's fine-to-coarse entries) are written by the second
 * of the level's two steps in the recursion, the one the transfer to the coarser level
 * follows; source records (the cells this level's coarse-to-fine entries read) by every
 * update.  psl_diad_apply_local_cell_update writes both.  Cells a transfer rewrites after
 * the update, and bounceback cells, are read from their slots; set_f, fill_equilibrium,
 * the outlet and a transfer into the level mark the records stale until the next update.
 * count: records of this level (0 before the first psl_diad_update_grid of its level list);
 * fresh: 1 when every one holds the last local update's values. */
int psl_diad_interface_records(const psl_diad* grid, int64_t* count, int* fresh);
/* Records whose values differ from the populations their cells hold now (0 right after a
 * local update of the level; a consistency check). */
int psl_diad_verify_interface_records(psl_diad* grid, int64_t* mismatches);
/* Restart state and health of an esotwist grid, as psl_grid_*_state / psl_grid_check_health:
 * the header carries the esotwist parity (esotwistFlipper) instead of a shifter, then the
 * storage words of f; loading marks the level's interface records stale. */
int psl_diad_state_bytes(const psl_diad* grid, int64_t* bytes);
int psl_diad_save_state(psl_diad* grid, void* host, int64_t bytes);
int psl_diad_load_state(psl_diad* grid, const void* host, int64_t bytes);
int psl_diad_check_health(psl_diad* grid, int64_t* bad_cells, double* rho_min, double* rho_max);
/* outletCellArray (include/types.h:171) and kMinNbrArray (include/types.h:161, n
 * entries) from host arrays; psl_diad_create_level sets them for the levels
 * whose outlet rules marked cells. */
int psl_diad_set_outlet(psl_diad* grid, const int* outlet_cells, int64_t n_outlet, const int* kmin);
/* applyNonReflectiveOutletZ (include/boundaryConditions/applyNonReflectiveOutlet.h:9-73):
 * for the nine c_z < 0 populations of every outlet cell,
 * f = invSqrt3 f(-k neighbour) + (1 - invSqrt3) f(cell) from the last update's
 * write maps, stored at the read maps of the next step; called before
 * psl_diad_apply_streaming. */
int psl_diad_apply_nonreflective_outlet_z(psl_diad* grid);
/* Moving walls (Marker.movingBounceback with the wall velocity of
 * getGivenRhoUxUyUz; M35Full/applyLocalCellUpdateCustomized.h:69-77, applyMovingBounceback
 * include/boundaryConditions/applyMovingBounceback.h:8-62): n distinct cells and
 * their velocities u3[3 n].  The cells must be flagged PSL_BOUNCEBACK; each
 * local update (and psl_diad_update_grid) then applies the moving-wall
 * bounceback to them.  n = 0 clears the list. */
int psl_diad_set_moving_bounceback(psl_diad* grid, const int* cells, int64_t n, const float* u3);
/* Forced-velocity cells (Marker.forcedVelocity with getForcing;
 * M35Full/applyLocalCellUpdateCustomized.h:80-86,114-121, M35Full/m35Main.cu:100-118, forcing
 * collision M35Full/applyCollisionCustomized.h:5-17): n distinct cells and their
 * target velocities ut3[3 n] (lattice units).  The cells must be flagged
 * PSL_BOUNCEBACK, so that the bulk update skips them; each local update (and
 * psl_diad_update_grid) then updates them with the force
 * g = rho (u_target - u) and the forcing collision.  A cell may not be both a
 * moving-wall and a forced cell.  n = 0 clears the list. */
int psl_diad_set_forced_velocity(psl_diad* grid, const int* cells, int64_t n, const float* ut3);
/* getSphereDrag(DIADGridStruct&) (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:132-198),
 * all three components in FP64: sum over BOUNCEBACK-flagged cells of
 * rho (u - u after bounceback), populations through the current write maps.
 * Moving-wall and forced cells are left out, as the reference's drag counts
 * neither its movingBounceback nor its forcedVelocity cells. */
int psl_diad_reduce_momentum_exchange(psl_diad* grid, double out[3]);
/* getTorque (M35Full/m35Main.cu:176-260) in FP64 over the forced cells, six
 * components: out[0..2] = sum of the forces g = rho (u_target - u), from the
 * populations the next update reads (as getTorque reads them), in newtons
 * (convertToPhysicalForce, include/cellFunctions.h:221-226, with the nominal
 * density rho_nominal_phys in kg/m3), out[3..5] = sum of r x g about the origin
 * in newton metres, r = (i res + ox, j res + oy, k res + oz) in millimetres.
 * out[5] is the reference's torque; its k < 5 and interface-cell exclusions
 * are the caller's choice of the forced-cell list. */
int psl_diad_reduce_forcing(psl_diad* grid, float rho_nominal_phys, double out[6]);

/* ------------------------------------------------- scalar transport */
/* ScalarTransportStruct (include/types.h:121) on a periodic-shift grid with FP32
 * distributions: 27 scalar populations per cell streamed by the grid's
 * shifter.  enable allocates them at zero (TArray.setValue(0)) with relaxation
 * time tauT = 3 D + 1/2. */
int psl_scalar_enable(psl_grid* grid, float tau_t);
/* getScalarTransportMarkers / getGivenT: a prescribed value per cell (N
 * floats, NaN where none), or NULL for none. */
int psl_scalar_set_given(psl_grid* grid, const float* given);
/* Scalar populations in logical order: host[q * N + cell]. */
int psl_scalar_get(psl_grid* grid, float* host);
int psl_scalar_set(psl_grid* grid, const float* host);
/* applyLocalCellUpdate(GridStruct&, ScalarTransportStruct&) (include/scalarTransportFunctions.h:10-103)
 * with applyScalarTransportCollision (lines 106-170): the fluid update of
 * psl_apply_local_cell_update plus the scalar, in one launch.  Streaming is
 * psl_apply_streaming. */
int psl_scalar_apply_local_cell_update(psl_grid* grid);
/* applyZeroGradientOutletScalarTransportBC (scalarTransport/scalarTransportMain.cu:95-135):
 * every non-bounceback cell of the last x layer takes the 27 scalar
 * populations of its -x neighbour; called between streaming and the update. */
int psl_scalar_apply_zero_gradient_outlet_x(psl_grid* grid);

/* ----------------------------------------------------------- output */
/* exportSectionCutPlotXY/ZY/ZX(GridStruct&) (include/plotter/exportSectionCutPlot.h:2-104):
 * plane 0 = XY at k = cut, 1 = ZY at i = cut, 2 = ZX at j = cut; rows x cols
 * pixels (dims[0] = rows, dims[1] = cols; host_out = NULL only reports them)
 * of 5 floats: p = (rho - 1) rho_nominal soundspeed^2, the horizontal,
 * vertical and normal velocity in m/s (u res / 1000 / dtPhys), and 1 for a
 * bounceback cell. */
int psl_section_cut(psl_grid* grid, int plane, int cut, float rho_nominal, float soundspeed, float* host_out,
                    int* dims);
/* exportSectionCutPlotXY/ZY/ZX(GridStruct&, ScalarTransportStruct&)
 * (include/plotter/exportSectionCutPlot.h:108-221): as psl_section_cut with a
 * sixth float per pixel, the scalar sum f T / sum f (getScalarTransport,
 * include/cellFunctions.h:228-235); needs psl_scalar_enable. */
int psl_scalar_section_cut(psl_grid* grid, int plane, int cut, float rho_nominal, float soundspeed, float* host_out,
                           int* dims);
/* The same over DIAD levels (exportSectionCutPlotGeneral(std::vector<DIADGridStruct>&), lines 377-589):
 * the plane at the finest level whose plane has fewer than max_pixels pixels
 * (dims3 = rows, cols, levels used; the reference uses 20 000 000), 6 floats
 * per pixel (the five above and the level); cut is an index on the finest
 * level.  Levels paint coarse to fine; where several finer cells fall on one
 * pixel the one with the largest cell index wins (the reference leaves that
 * to the order of its threads). */
int psl_diad_section_cut(psl_diad* const* levels, int nlevels, int plane, int cut, float rho_nominal,
                         float soundspeed, int64_t max_pixels, float* host_out, int* dims3);
/* getFlowReportXY/ZY/ZX (include/flowReportFunctions.h:2-263): the DIAD plane cut
 * clipped to bounds6 = (xmin, ymin, zmin, xmax, ymax, zmax) (physical units),
 * then over its pixels outside walls the mean ux, uy, uz and rho in lattice
 * units and the area in mm^2: out5 = ux, uy, uz, rho, area (FP64 sums). */
int psl_diad_flow_report(psl_diad* const* levels, int nlevels, int plane, int cut, const float* bounds6,
                         double* out5);
/* /dev/shm/sim_data.bin format of the reference's plotters (exportSectionCutPlot.h:64-86):
 * int {plot_number, rows, cols, components}, then the floats. */
int psl_write_sim_data(const char* path, int plot_number, int rows, int cols, int components, const float* data);
/* exportSectionCutPlotToiletPaperZ(GridStruct&) (include/plotter/exportSectionCutPlot.h:674-787):
 * the cylinder of radius r (mm) about the z axis unrolled; dims[0] =
 * (int)(2 pi r / res) rows along the circumference (row v at angle v res / r,
 * the cell of getIJKCellIndexFromXYZ), dims[1] = cellCountZ columns; 5 floats
 * per pixel (p, uz, tangential and radial velocity in m/s, bounceback
 * marker).  host_out = NULL only reports dims.  A radius whose circle leaves
 * the grid is rejected (the reference reads outside its arrays there). */
int psl_cylinder_cut(psl_grid* grid, float r, float rho_nominal, float soundspeed, float* host_out, int* dims);
/* The same over DIAD levels (exportSectionCutPlotToiletPaperZ(std::vector<DIADGridStruct>&),
 * lines 789-993): the finest level whose unrolled plane has fewer than
 * max_pixels pixels sets the pixel size (dims3 = rows, cols, levels used);
 * cells within 1.5 half-widths of the cylinder paint their footprint at their
 * polar angle, rows wrapping round the circumference, coarse to fine, the
 * largest cell index winning where cells of one level share a pixel (the
 * reference leaves that to the order of its writes); 6 floats per pixel (the
 * five above and the level). */
int psl_diad_cylinder_cut(psl_diad* const* levels, int nlevels, float r, float rho_nominal, float soundspeed,
                          int64_t max_pixels, float* host_out, int* dims3);
/* export3DPlot(GridStruct&) (include/plotter/exportSectionCutPlot.h:591-672): every
 * cell in (i, j, k) order with k fastest, 5 floats each (p, ux, uy, uz in m/s
 * as in psl_section_cut, bounceback marker); host_out holds 5 N floats. */
int psl_volume_export(psl_grid* grid, float rho_nominal, float soundspeed, float* host_out);
/* Its file: int plot number, nx, ny, nz, components, then the floats. */
int psl_write_volume_data(const char* path, int plot_number, int nx, int ny, int nz, int components,
                          const float* data);
/* historyData.bin (exportSectionCutPlot.h:994-1001): int count, then count floats. */
int psl_write_history(const char* path, const float* values, int count);

/* ------------------------------------------------------------- geometry */
/* readSTL (include/STLFunctions.h:3-89). */
int psl_stl_read(const char* path, psl_stl** out);
/* STL from n triangles of 9 floats (ax ay az bx by bz cx cy cz). */
int psl_stl_from_triangles(const float* tri9, int n, psl_stl** out);
int psl_stl_destroy(psl_stl* stl);
/* triangleCount and bounds xmin ymin zmin xmax ymax zmax. */
int psl_stl_info(const psl_stl* stl, int* triangle_count, float* bounds6);
int psl_stl_set_origin(psl_stl* stl, float ox, float oy, float oz);
int psl_stl_get_triangles(const psl_stl* stl, float* tri9);
/* rotateSTLAlongX / Y / Z (include/STLFunctions.h:698-809): rotates every
 * vertex by `radians` about axis 0 (x), 1 (y) or 2 (z) through the STL's own
 * origin (right-handed), in FP32. */
int psl_stl_rotate(psl_stl* stl, int axis, float radians);
/* ApplyMarkersInsideSTL (include/STLFunctions.h:482-696) on a DIAD grid's cells:
 * out[c] = 1 for a cell inside the STL, 0 outside (n host bytes).  The
 * M35 drivers re-mark their forced cells this way after each rotation. */
int psl_diad_markers_inside_stl(const psl_diad* grid, const psl_stl* stl, uint8_t* out);
/* checkSTLEdges (include/STLFunctions.h:92-182): triangles whose edges are not shared
 * exactly twice; per-edge counts (3 per triangle) when counts != NULL. */
int psl_stl_check_edges(const psl_stl* stl, int* bad_triangles, int* counts);

/* BoolArrayType of cellCount entries on the device. */
int psl_mask_create(int64_t n, int value, psl_mask** out);
int psl_mask_destroy(psl_mask* m);
int psl_mask_get(const psl_mask* m, uint8_t* host);
int psl_mask_set(psl_mask* m, const uint8_t* host);
/* multiplyBoolArrays / sumBoolArrays / invertBoolArray (include/STLFunctions.h:811-846). */
int psl_mask_and(const psl_mask* a, const psl_mask* b, psl_mask* out);
int psl_mask_or(const psl_mask* a, const psl_mask* b, psl_mask* out);
int psl_mask_not(psl_mask* m);
/* applyMarkersInsideSTL (include/STLFunctions.h:277-479). */
int psl_markers_inside_stl(psl_mask* mask, const psl_stl* stl, int inside_value, const psl_info* info);
/* Set the kind of every cell with mask != 0 to PSL_BOUNCEBACK (Grid.bouncebackMarkerArray). */
int psl_grid_apply_bounceback_mask(psl_grid* grid, const psl_mask* mask);

#ifdef __cplusplus
}
#endif
#endif /* PSL_H */
