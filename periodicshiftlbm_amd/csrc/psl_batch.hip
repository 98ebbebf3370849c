// Batched periodic-shift steps over many independent grids of uneven size.
//
// No reference counterpart: the reference advances one grid at a time
// (flowAroundSphere/flowAroundSphereMainBatch.cu loops over Reynolds numbers
// sequentially).  Here every grid keeps its own psl_grid (flags, boundary
// list, given values, storage) and a batch advances all of them together:
//
//   k_batch_advance   applyStreaming of every grid on the device: the shift
//                     update and the per-population bases of the step, written
//                     into each grid's StepParams;
//   k_batch_bulk      one launch over the bulk tiles of all grids (tile ->
//                     grid and tile -> block tables): 256 or 512 cells per
//                     workgroup (one or two cells per thread), the grids'
//                     tiles in turn or interleaved across grids;
//   k_batch_boundary  one launch over the concatenated boundary-list blocks;
//   k_batch_fused     both in one launch, each grid's boundary workgroups
//                     interleaved with its bulk blocks (a tuning variant).
//
// The per-cell work is bulk_update / boundary_update, the same functions on
// the same StepParams / BoundaryParams as the single-grid kernels, so every
// grid's result is bit-identical to stepping it alone.  Steps are captured
// in hipGraphs, one of a single step and one of steps_per_graph (default
// 1000) consecutive steps, and replayed: 5000 steps are five graph launches
// whatever the number of grids.  The block order of the grids (as given, or largest
// first) and the kernel variant are chosen by psl_batch_autotune, and with
// them the cells per thread and the tile order by psl_batch_autotune_layouts.
#include "psl_internal.h"
#include "psl_stream.h"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace psl {

// applyStreaming of every grid (include/applyStreaming.h:11-31) and the
// per-population stream bases of the step.
template <class St>
__global__ void k_batch_advance(StepParams* __restrict__ params, const void* const* __restrict__ fbase,
                                const int* __restrict__ delta, int* __restrict__ shift, int nd) {
  typedef typename St::T T;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nd * 27) return;
  const int d = t / 27, q = t - d * 27;
  StepParams& P = params[d];
  const int64_t N = P.n;
  int64_t s = shift[t];
  s -= delta[t];
  // reduced modulo N (also on grids thinner than two wraps allow)
  s %= N;
  if (s < 0) s += N;
  shift[t] = (int)s;
  T* row = (T*)fbase[d] + (size_t)q * P.n + s;
  P.lo[q] = row;
  P.hi[q] = row - P.n;
  P.thr[q] = P.n - (uint32_t)s;
}

// One tile of CPT consecutive 256-cell blocks of grid tile_domain[t], its
// tile index tile_index[t] within the grid; each thread updates one cell of
// every block in turn.
template <class St, int TUNE, int CPT>
__global__ __launch_bounds__(256) void k_batch_bulk(const StepParams* __restrict__ params,
                                                    const uint32_t* __restrict__ tile_domain,
                                                    const uint32_t* __restrict__ tile_index, uint32_t remap) {
  const uint32_t t = remap ? xcd_block_chunked<64>(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t d = tile_domain[t];
  const uint32_t first = tile_index[t] * CPT;
  for (int k = 0; k < CPT; ++k) bulk_update<St, true, true, TUNE>(params[d], first + k);
}

template <class St>
__global__ __launch_bounds__(256) void k_batch_boundary(const StepParams* __restrict__ params,
                                                        const BoundaryParams* __restrict__ bparams,
                                                        const uint32_t* __restrict__ bblock_domain,
                                                        const uint32_t* __restrict__ bblock_first) {
  const uint32_t d = bblock_domain[blockIdx.x];
  boundary_update<St>(params[d], bparams[d], (blockIdx.x - bblock_first[d]) * 256u + threadIdx.x);
}

// One workgroup of the fused batch launch: entry `e` of grid `d`'s fused
// order (boundary-list ranges interleaved with the bulk blocks they border).
struct BatchEntry {
  uint32_t d;
  FusedEntry e;
  uint32_t pad;
};

// Held at 6 waves per SIMD: the boundary branch then spills 30 VGPRs instead
// of 46 at 7 waves, and the launch is ~2 % faster on the 64-domain batch.
template <class St, int TUNE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_batch_fused(
    const StepParams* __restrict__ params, const BoundaryParams* __restrict__ bparams,
    const BatchEntry* __restrict__ order, uint32_t remap) {
  const uint32_t w = remap ? xcd_block_chunked<64>(blockIdx.x, gridDim.x) : blockIdx.x;
  const BatchEntry x = order[w];
  if (x.e.b != FUSED_BULK) {
    BoundaryParams b = bparams[x.d];
    b.n = x.e.b;
    boundary_update<St>(params[x.d], b, x.e.a + threadIdx.x);
    return;
  }
  bulk_update<St, true, true, TUNE>(params[x.d], x.e.a);
}

}  // namespace psl

using namespace psl;

struct psl_batch {
  std::vector<psl_grid*> grids;
  int storage = PSL_F32;
  int tuning = PSL_TUNE_XCD_REMAP;
  int order = 0;  // 0: as given, 1: largest grid first
  int cpt = 1;         // cells per thread of the two-launch bulk kernel (1 or 2)
  int tile_order = 0;  // 0: each grid's tiles in turn, 1: interleaved across grids
  hipStream_t stream = nullptr;
  StepParams* d_params = nullptr;
  BoundaryParams* d_bparams = nullptr;
  const void** d_fbase = nullptr;
  int* d_delta = nullptr;
  int* d_shift = nullptr;
  uint32_t* d_block_domain = nullptr;  // bulk tile -> grid
  uint32_t* d_block_tile = nullptr;    // bulk tile -> tile index within its grid
  uint32_t* d_bblock_domain = nullptr;
  uint32_t* d_bblock_first = nullptr;
  BatchEntry* d_order = nullptr;
  uint32_t total_blocks = 0, total_bblocks = 0, total_bnd = 0, total_fused = 0;
  int64_t total_cells = 0;
  hipGraph_t graph = nullptr;  // one step
  hipGraphExec_t exec = nullptr;
  int steps_per_graph = 1000;  // steps captured in graph_k
  hipGraph_t graph_k = nullptr;
  hipGraphExec_t exec_k = nullptr;
  // what the captured graph was built from
  std::vector<StepParams> h_params;
  std::vector<BoundaryParams> h_bparams;
  std::vector<const void*> h_fbase;
};

namespace {

void free_graphs(psl_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->exec_k) (void)hipGraphExecDestroy(b->exec_k);
  if (b->graph_k) (void)hipGraphDestroy(b->graph_k);
  b->exec = nullptr;
  b->graph = nullptr;
  b->exec_k = nullptr;
  b->graph_k = nullptr;
}

void free_batch_device(psl_batch* b) {
  free_graphs(b);
  void* ptrs[] = {b->d_params,       b->d_bparams,     b->d_fbase,         b->d_delta,          b->d_shift,
                  b->d_block_domain, b->d_block_tile, b->d_bblock_domain, b->d_bblock_first, b->d_order};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  b->d_params = nullptr;
  b->d_bparams = nullptr;
  b->d_fbase = nullptr;
  b->d_delta = nullptr;
  b->d_shift = nullptr;
  b->d_block_domain = nullptr;
  b->d_block_tile = nullptr;
  b->d_bblock_domain = nullptr;
  b->d_bblock_first = nullptr;
  b->d_order = nullptr;
}

template <class St>
int enqueue_step(psl_batch* b) {
  const int nd = (int)b->grids.size();
  k_batch_advance<St><<<(nd * 27 + 255) / 256, 256, 0, b->stream>>>(b->d_params, b->d_fbase, b->d_delta,
                                                                     b->d_shift, nd);
  const uint32_t remap = (b->tuning & PSL_TUNE_XCD_REMAP) ? 1u : 0u;
  const int nt = b->tuning & (TUNE_NT_LOAD | TUNE_NT_STORE);
  if ((b->tuning & PSL_TUNE_FUSED_BOUNDARY) && b->total_bnd) {
#define PSL_FUSED(V) \
  k_batch_fused<St, V><<<b->total_fused, 256, 0, b->stream>>>(b->d_params, b->d_bparams, b->d_order, remap)
    switch (nt) {
      case 0: PSL_FUSED(0); break;
      case 1: PSL_FUSED(1); break;
      case 2: PSL_FUSED(2); break;
      default: PSL_FUSED(3); break;
    }
#undef PSL_FUSED
  } else {
#define PSL_BULK(V, C)                                                                                          \
  k_batch_bulk<St, V, C><<<b->total_blocks, 256, 0, b->stream>>>(b->d_params, b->d_block_domain, b->d_block_tile, \
                                                                 remap)
    if (b->cpt == 2) {
      switch (nt) {
        case 0: PSL_BULK(0, 2); break;
        case 1: PSL_BULK(1, 2); break;
        case 2: PSL_BULK(2, 2); break;
        default: PSL_BULK(3, 2); break;
      }
    } else {
      switch (nt) {
        case 0: PSL_BULK(0, 1); break;
        case 1: PSL_BULK(1, 1); break;
        case 2: PSL_BULK(2, 1); break;
        default: PSL_BULK(3, 1); break;
      }
    }
#undef PSL_BULK
    if (b->total_bnd)
      k_batch_boundary<St><<<b->total_bblocks, 256, 0, b->stream>>>(b->d_params, b->d_bparams, b->d_bblock_domain,
                                                                     b->d_bblock_first);
  }
  PSL_LAUNCH_CHECK();
  return PSL_OK;
}

int enqueue_dispatch(psl_batch* b) {
  switch (b->storage) {
    case PSL_F32: return enqueue_step<StF32>(b);
    case PSL_F16DEV: return enqueue_step<StF16>(b);
    default: return enqueue_step<StBF16>(b);
  }
}

template <class T>
int upload(T** dst, const std::vector<T>& src) {
  PSL_HIP_TRY(hipMalloc((void**)dst, std::max<size_t>(1, src.size()) * sizeof(T)));
  if (!src.empty()) PSL_HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return PSL_OK;
}

// Capture `steps` consecutive steps into one graph (the kernels read all
// state through device tables, so replaying it advances every grid `steps`
// times) and instantiate it.
int capture_steps(psl_batch* b, int steps, hipGraph_t* graph_out, hipGraphExec_t* exec_out) {
  hipStream_t cap = nullptr;
  PSL_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipStream_t saved = b->stream;
  b->stream = cap;
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  int rc = PSL_OK;
  for (int s = 0; e == hipSuccess && rc == PSL_OK && s < steps; ++s) rc = enqueue_dispatch(b);
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(cap, &graph);
  b->stream = saved;
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess || e2 != hipSuccess || rc != PSL_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    set_error("batch graph capture failed");
    return rc ? rc : PSL_E_HIP;
  }
  const hipError_t e3 = hipGraphInstantiate(exec_out, graph, nullptr, nullptr, 0);
  if (e3 != hipSuccess) {
    (void)hipGraphDestroy(graph);
    set_error(std::string("batch graph instantiation failed: ") + hipGetErrorString(e3));
    return PSL_E_HIP;
  }
  *graph_out = graph;
  return PSL_OK;
}

// Host-side StepParams / BoundaryParams of a grid (the stream bases are
// filled in on the device each step).
StepParams step_params_of(const psl_grid* g) {
  StepParams p;
  memset(&p, 0, sizeof(p));
  p.n = (uint32_t)g->info.cellCount;
  const float tau = 3.f * g->info.nu + 0.5f;
  p.tau = tau;
  p.omega = 1.f / tau;
  p.cs = g->cs;
  p.cs_cell = g->cs_cell;
  p.flags = g->flags;
  return p;
}

BoundaryParams boundary_params_of(const psl_grid* g) {
  BoundaryParams q;
  memset(&q, 0, sizeof(q));
  q.list = g->bc_list;
  q.n = g->n_bc;
  q.nx = g->info.cellCountX;
  q.ny = g->info.cellCountY;
  q.nz = g->info.cellCountZ;
  q.bc = g->bc_dev;
  return q;
}

// Build the device tables for the current order and capture one step.
int build_batch(psl_batch* b) {
  free_batch_device(b);
  const int nd = (int)b->grids.size();
  std::vector<int> order(nd);
  std::iota(order.begin(), order.end(), 0);
  if (b->order == 1)
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->grids[x]->info.cellCount > b->grids[y]->info.cellCount; });
  std::vector<StepParams> params(nd);
  std::vector<BoundaryParams> bparams(nd);
  std::vector<const void*> fbase(nd);
  std::vector<int> delta((size_t)nd * 27);
  std::vector<uint32_t> block_domain, block_tile, bblock_domain, bblock_first(nd);
  b->total_cells = 0;
  b->total_bnd = 0;
  // bulk tiles of 256 * cpt cells: each grid's in turn, or round-robin
  // across the grids (tile t of every grid that has one, then t + 1)
  const uint64_t tile_cells = 256ull * (uint64_t)b->cpt;
  std::vector<uint32_t> ntile(nd);
  uint32_t max_tiles = 0;
  for (int d = 0; d < nd; ++d) {
    ntile[d] = (uint32_t)((b->grids[d]->info.cellCount + tile_cells - 1) / tile_cells);
    max_tiles = std::max(max_tiles, ntile[d]);
  }
  if (b->tile_order == 1) {
    for (uint32_t t = 0; t < max_tiles; ++t)
      for (int d : order)
        if (t < ntile[d]) {
          block_domain.push_back((uint32_t)d);
          block_tile.push_back(t);
        }
  } else {
    for (int d : order)
      for (uint32_t t = 0; t < ntile[d]; ++t) {
        block_domain.push_back((uint32_t)d);
        block_tile.push_back(t);
      }
  }
  for (int d : order) {
    const psl_grid* g = b->grids[d];
    const uint32_t nbb = (g->n_bc + 255u) / 256u;
    bblock_first[d] = (uint32_t)bblock_domain.size();
    bblock_domain.insert(bblock_domain.end(), nbb, (uint32_t)d);
  }
  // the fused order: every grid's own interleaving, grids in launch order
  std::vector<BatchEntry> fused;
  for (int d : order) {
    const psl_grid* g = b->grids[d];
    if (g->fused_host.empty()) {
      const uint32_t nbk = (uint32_t)((g->info.cellCount + 255) / 256);
      for (uint32_t k = 0; k < nbk; ++k) fused.push_back({(uint32_t)d, {k, FUSED_BULK}, 0u});
    } else {
      for (const FusedEntry& e : g->fused_host) fused.push_back({(uint32_t)d, e, 0u});
    }
  }
  for (int d = 0; d < nd; ++d) {
    const psl_grid* g = b->grids[d];
    params[d] = step_params_of(g);
    bparams[d] = boundary_params_of(g);
    fbase[d] = g->f;
    for (int q = 0; q < 27; ++q)
      delta[(size_t)d * 27 + q] = CX[q] + g->info.cellCountX * CY[q] + g->info.cellCountX * g->info.cellCountY * CZ[q];
    b->total_bnd += g->n_bc;
    b->total_cells += g->info.cellCount;
  }
  b->total_blocks = (uint32_t)block_domain.size();
  b->total_bblocks = (uint32_t)bblock_domain.size();
  b->total_fused = (uint32_t)fused.size();
  b->h_params = params;
  b->h_bparams = bparams;
  b->h_fbase = fbase;
  int rc = upload(&b->d_params, params);
  if (!rc) rc = upload(&b->d_bparams, bparams);
  if (!rc) rc = upload(&b->d_fbase, fbase);
  if (!rc) rc = upload(&b->d_delta, delta);
  if (!rc) rc = upload(&b->d_block_domain, block_domain);
  if (!rc) rc = upload(&b->d_block_tile, block_tile);
  if (!rc) rc = upload(&b->d_bblock_domain, bblock_domain);
  if (!rc) rc = upload(&b->d_bblock_first, bblock_first);
  if (!rc) rc = upload(&b->d_order, fused);
  if (!rc) rc = upload(&b->d_shift, std::vector<int>((size_t)nd * 27, 0));
  if (rc) return rc;
  return capture_steps(b, 1, &b->graph, &b->exec);
}

// True when a grid changed since the graph was captured (new flags, boundary
// list, given-value table, viscosity or Smagorinsky constant).
bool batch_stale(const psl_batch* b) {
  for (size_t d = 0; d < b->grids.size(); ++d) {
    const psl_grid* g = b->grids[d];
    const StepParams p = step_params_of(g);
    const BoundaryParams q = boundary_params_of(g);
    const StepParams& P = b->h_params[d];
    const BoundaryParams& Q = b->h_bparams[d];
    if (p.n != P.n || p.tau != P.tau || p.cs != P.cs || p.cs_cell != P.cs_cell || p.flags != P.flags ||
        q.list != Q.list || q.n != Q.n || q.bc != Q.bc || g->f != b->h_fbase[d])
      return true;
  }
  return false;
}

// A grid the batch kernels can advance: flags set (the boundary list is
// built from them), no learned closure, no periodic copies, no sub-grid.
const char* unbatchable(const psl_grid* g) {
  if (!g->flags) return "batch grids need flags (psl_grid_set_flags)";
  if (g->closure) return "batch grids cannot use the learned closure";
  if (g->periodic_copies) return "batch grids cannot use periodic copies";
  if (g->info.iSubgridEnd - g->info.iSubgridStart > 4) return "batch grids cannot carry a sub-grid";
  return nullptr;
}

int upload_shifts(psl_batch* b) {
  const int nd = (int)b->grids.size();
  std::vector<int> sh((size_t)nd * 27);
  for (int d = 0; d < nd; ++d)
    for (int q = 0; q < 27; ++q) sh[(size_t)d * 27 + q] = b->grids[d]->shift[q];
  PSL_HIP_TRY(hipMemcpyAsync(b->d_shift, sh.data(), sh.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  PSL_HIP_TRY(hipStreamSynchronize(b->stream));
  return PSL_OK;
}

}  // namespace

extern "C" {

int psl_batch_create(psl_grid* const* grids, int count, psl_batch** out) {
  PSL_CHECK_ARG(grids && count > 0 && out, "bad grid list");
  psl_batch* b = new psl_batch();
  for (int d = 0; d < count; ++d) {
    psl_grid* g = grids[d];
    const char* why = !g ? "null grid"
                         : g->storage != grids[0]->storage ? "batch grids must share the storage" : unbatchable(g);
    if (why) {
      delete b;
      set_error(why);
      return PSL_E_ARG;
    }
    b->grids.push_back(g);
  }
  b->storage = grids[0]->storage;
  b->stream = grids[0]->stream;
  int rc = build_batch(b);
  if (rc) {
    free_batch_device(b);
    delete b;
    return rc;
  }
  *out = b;
  return PSL_OK;
}

int psl_batch_destroy(psl_batch* b) {
  if (!b) return PSL_OK;
  (void)hipDeviceSynchronize();
  free_batch_device(b);
  delete b;
  return PSL_OK;
}

// nsteps full steps (applyStreaming + applyLocalCellUpdate) of every grid.
// The grids' host shifters follow, so single-grid calls remain valid after.
int psl_batch_step(psl_batch* b, int nsteps) {
  PSL_CHECK_ARG(b && nsteps >= 0, "bad batch step");
  if (nsteps == 0) return PSL_OK;
  // a closure, periodic copies or a sub-grid set on a grid after the batch
  // was created are refused here, whatever the captured graph holds
  for (const psl_grid* g : b->grids) {
    const char* why = unbatchable(g);
    if (why) {
      set_error(why);
      return PSL_E_ARG;
    }
  }
  if (batch_stale(b)) {
    int rc = build_batch(b);
    if (rc) return rc;
  }
  int rc = upload_shifts(b);
  if (rc) return rc;
  // whole graphs of steps_per_graph steps, then single steps
  const int k = b->steps_per_graph;
  if (k > 1 && nsteps >= k && !b->exec_k) {
    rc = capture_steps(b, k, &b->graph_k, &b->exec_k);
    if (rc) return rc;
  }
  int left = nsteps;
  if (k > 1)
    for (; left >= k; left -= k) PSL_HIP_TRY(hipGraphLaunch(b->exec_k, b->stream));
  for (; left > 0; --left) PSL_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  for (psl_grid* g : b->grids)
    for (int s = 0; s < nsteps; ++s) psl_apply_streaming(g);
  return PSL_OK;
}

int psl_batch_set_tuning(psl_batch* b, int tuning, int order) {
  PSL_CHECK_ARG(b, "null batch");
  PSL_CHECK_ARG(tuning >= 0 && tuning < PSL_TUNE_VARIANTS && (order == 0 || order == 1), "bad tuning");
  PSL_HIP_TRY(hipStreamSynchronize(b->stream));
  b->tuning = tuning;
  b->order = order;
  return build_batch(b);
}

int psl_batch_set_layout(psl_batch* b, int cells_per_thread, int tile_order) {
  PSL_CHECK_ARG(b, "null batch");
  PSL_CHECK_ARG(cells_per_thread == 1 || cells_per_thread == 2, "cells per thread must be 1 or 2");
  PSL_CHECK_ARG(tile_order == 0 || tile_order == 1, "tile order must be 0 (grids in turn) or 1 (interleaved)");
  PSL_HIP_TRY(hipStreamSynchronize(b->stream));
  b->cpt = cells_per_thread;
  b->tile_order = tile_order;
  return build_batch(b);
}

int psl_batch_get_layout(const psl_batch* b, int* cells_per_thread, int* tile_order) {
  PSL_CHECK_ARG(b, "null batch");
  if (cells_per_thread) *cells_per_thread = b->cpt;
  if (tile_order) *tile_order = b->tile_order;
  return PSL_OK;
}

int psl_batch_set_steps_per_graph(psl_batch* b, int steps) {
  PSL_CHECK_ARG(b && steps >= 1 && steps <= PSL_BATCH_MAX_STEPS_PER_GRAPH, "steps per graph out of range");
  if (steps == b->steps_per_graph) return PSL_OK;
  PSL_HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->exec_k) (void)hipGraphExecDestroy(b->exec_k);
  if (b->graph_k) (void)hipGraphDestroy(b->graph_k);
  b->exec_k = nullptr;
  b->graph_k = nullptr;
  b->steps_per_graph = steps;
  return PSL_OK;
}

int psl_batch_get_steps_per_graph(const psl_batch* b, int* steps) {
  PSL_CHECK_ARG(b && steps, "null argument");
  *steps = b->steps_per_graph;
  return PSL_OK;
}

int psl_batch_get_tuning(const psl_batch* b, int* tuning, int* order) {
  PSL_CHECK_ARG(b, "null batch");
  if (tuning) *tuning = b->tuning;
  if (order) *order = b->order;
  return PSL_OK;
}

// Time every (variant, grid order) configuration for `steps` graph steps each
// after one warm-up step (the grids advance) and keep the fastest.
// times_ms[2 * PSL_TUNE_VARIANTS] receives ms per step at order * 16 + variant.
int psl_batch_autotune(psl_batch* b, int steps, int* best_tuning, int* best_order, double* times_ms) {
  PSL_CHECK_ARG(b && steps > 0, "bad autotune arguments");
  double best = 1e300;
  int bt = b->tuning, bo = b->order;
  hipEvent_t e0, e1;
  PSL_HIP_TRY(hipEventCreate(&e0));
  PSL_HIP_TRY(hipEventCreate(&e1));
  for (int order = 0; order < 2; ++order)
    for (int t = 0; t < PSL_TUNE_VARIANTS; ++t) {
      int rc = psl_batch_set_tuning(b, t, order);
      if (!rc) rc = psl_batch_step(b, 1);  // warm
      if (!rc && hipEventRecord(e0, b->stream) != hipSuccess) rc = PSL_E_HIP;
      if (!rc) rc = psl_batch_step(b, steps);
      if (!rc && hipEventRecord(e1, b->stream) != hipSuccess) rc = PSL_E_HIP;
      float ms = 0.f;
      if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = PSL_E_HIP;
      if (rc) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
      }
      const double per = ms / steps;
      if (times_ms) times_ms[order * PSL_TUNE_VARIANTS + t] = per;
      if (per < best) {
        best = per;
        bt = t;
        bo = order;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best_tuning) *best_tuning = bt;
  if (best_order) *best_order = bo;
  return psl_batch_set_tuning(b, bt, bo);
}

// The autotune above at every layout: cells per thread 1 and 2 and both tile
// orders for the two-launch variants (the fused variants have their own
// workgroup order and one cell per thread, so they are timed at layout 1/0
// only).  times_ms[PSL_BATCH_LAYOUT_CONFIGS]: ms per step at
// ((tile_order * 2 + cpt - 1) * 2 + order) * PSL_TUNE_VARIANTS + variant, -1
// where a configuration is not run.
int psl_batch_autotune_layouts(psl_batch* b, int steps, int best[4], double* times_ms) {
  PSL_CHECK_ARG(b && steps > 0 && best, "bad autotune arguments");
  double bt = 1e300;
  int bv = b->tuning, bo = b->order, bc = b->cpt, btile = b->tile_order;
  if (times_ms)
    for (int i = 0; i < PSL_BATCH_LAYOUT_CONFIGS; ++i) times_ms[i] = -1.0;
  double part[2 * PSL_TUNE_VARIANTS];
  for (int tile = 0; tile < 2; ++tile)
    for (int c = 1; c <= 2; ++c) {
      int rc = psl_batch_set_layout(b, c, tile);
      if (!rc) rc = psl_batch_autotune(b, steps, nullptr, nullptr, part);
      if (rc) return rc;
      for (int order = 0; order < 2; ++order)
        for (int v = 0; v < PSL_TUNE_VARIANTS; ++v) {
          const bool fused = (v & PSL_TUNE_FUSED_BOUNDARY) != 0;
          if (fused && (c != 1 || tile != 0)) continue;  // the same launch as at layout 1/0
          const double per = part[order * PSL_TUNE_VARIANTS + v];
          if (times_ms) times_ms[((tile * 2 + c - 1) * 2 + order) * PSL_TUNE_VARIANTS + v] = per;
          if (per < bt) {
            bt = per;
            bv = v;
            bo = order;
            bc = c;
            btile = tile;
          }
        }
    }
  best[0] = bv;
  best[1] = bo;
  best[2] = bc;
  best[3] = btile;
  PSL_HIP_TRY(hipStreamSynchronize(b->stream));
  b->tuning = bv;
  b->order = bo;
  b->cpt = bc;
  b->tile_order = btile;
  return build_batch(b);
}

int psl_batch_census(const psl_batch* b, int64_t* cells, int* blocks, int* boundary_cells) {
  PSL_CHECK_ARG(b, "null batch");
  if (cells) *cells = b->total_cells;
  if (blocks) *blocks = (int)b->total_blocks;
  if (boundary_cells) *boundary_cells = (int)b->total_bnd;
  return PSL_OK;
}

}  // extern "C"
