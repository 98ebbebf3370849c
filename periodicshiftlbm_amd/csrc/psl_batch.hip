// Batched periodic-shift steps over many independent grids of uneven size.
//
// No reference counterpart: the reference advances one grid at a time
// (flowAroundSphere/flowAroundSphereMainBatch.cu loops over Reynolds numbers
// sequentially).  Here every grid keeps its own psl_grid (flags, boundary
// list, given values, storage) and a batch advances all of them together:
//
//   k_batch_advance   applyStreaming of every grid on the device: the shift
//                     update and the per-population base / threshold tables;
//   k_batch_bulk      one launch over the concatenated blocks of all grids
//                     (block -> grid table), the same per-cell path as k_ps_bulk;
//   k_batch_boundary  one launch over the concatenated box-face lists.
//
// One step is captured once in a hipGraph and replayed, so a step costs one
// graph launch whatever the number of grids.  The block order of the grids
// (as given, or largest first) and the bulk-kernel variant are chosen by
// psl_batch_autotune.
#include "psl_internal.h"
#include "psl_stream.h"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace psl {

struct DomainDesc {
  char* f;
  const uint8_t* flags;
  const uint32_t* bc_list;
  uint32_t n, n_bc;
  int nx, ny, nz;
  uint32_t block_first;
  float omega, tau, cs;
  int delta[27];  // shift decrement per step: c_x + Nx c_y + Nx Ny c_z
  const float4* bc;  // the grid's given-value slots (device)
};

// Per-grid tables of the current step (written by k_batch_advance).
struct DomainTable {
  const void* lo[27];
  const void* hi[27];
  uint32_t thr[27];
  uint32_t n;
};

// applyStreaming of every grid (include/applyStreaming.h:11-31) and the
// per-population stream bases of the step.
template <class St>
__global__ void k_batch_advance(const DomainDesc* __restrict__ desc, int* __restrict__ shift,
                                DomainTable* __restrict__ table, int nd) {
  typedef typename St::T T;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nd * 27) return;
  const int d = t / 27, q = t - d * 27;
  const DomainDesc& D = desc[d];
  const int64_t N = D.n;
  int64_t s = shift[t];
  s -= D.delta[q];
  if (s < 0) s += N;
  if (s < 0) s += N;
  if (s >= N) s -= N;
  if (s >= N) s -= N;
  shift[t] = (int)s;
  T* row = (T*)D.f + (size_t)q * D.n + s;
  table[d].lo[q] = row;
  table[d].hi[q] = row - D.n;
  table[d].thr[q] = D.n - (uint32_t)s;
  if (q == 0) table[d].n = D.n;
}

template <class St, int TUNE>
__global__ __launch_bounds__(256) void k_batch_bulk(const DomainDesc* __restrict__ desc,
                                                    const DomainTable* __restrict__ table,
                                                    const uint32_t* __restrict__ block_domain, uint32_t xcd_remap) {
  const uint32_t b = xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t d = block_domain[b];
  const DomainDesc& D = desc[d];
  const DomainTable& T = table[d];
  const uint32_t block_first = (b - D.block_first) * 256u;
  const uint32_t cell = block_first + threadIdx.x;
  const uint32_t w_first = wave_first_cell(block_first);
  if (w_first >= D.n) return;
  const bool straddle = wave_straddles(T, w_first);
  if (cell >= D.n) return;
  const int kind = D.flags ? (D.flags[cell] & 7) : KIND_FLUID;
  if (kind > KIND_BOUNCEBACK) return;
  float v[27];
  load27<St, TUNE>(T, w_first, straddle, cell, v);
  if (kind == KIND_BOUNCEBACK) {
    bounceback(v);
  } else {
    Relax rx{D.omega, D.tau, D.cs};
    collide<St::W_OFF>(v, rx, D.cs);
  }
  store27<St, TUNE>(T, w_first, straddle, cell, v);
}

template <class St>
__global__ __launch_bounds__(256) void k_batch_boundary(const DomainDesc* __restrict__ desc,
                                                        const DomainTable* __restrict__ table,
                                                        const uint32_t* __restrict__ bnd_domain,
                                                        const uint32_t* __restrict__ bnd_cell, uint32_t total) {
  typedef typename St::T T;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= total) return;
  const uint32_t d = bnd_domain[t];
  const uint32_t cell = bnd_cell[t];
  const DomainDesc& D = desc[d];
  const DomainTable& Tb = table[d];
  const int fl = D.flags[cell];
  const int kind = fl & 7;
  const uint32_t nxy = (uint32_t)D.nx * (uint32_t)D.ny;
  const int k = (int)(cell / nxy);
  const uint32_t rem = cell - (uint32_t)k * nxy;
  const int j = (int)(rem / (uint32_t)D.nx);
  const int i = (int)(rem - (uint32_t)j * (uint32_t)D.nx);
  float v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const T* base = (const T*)(cell >= Tb.thr[q] ? Tb.hi[q] : Tb.lo[q]);
    v[q] = St::load(base[cell]);
    if (St::W_OFF) v[q] += weight(q);
  }
  const int code = outer_normal_code(i, j, k, D.nx, D.ny, D.nz, fl);
  const float4 g = D.bc[(fl >> 6) & 3];
  boundary_cell(v, kind, code, g.x, g.y, g.z, g.w);
  Relax rx{D.omega, D.tau, D.cs};
  collide<false>(v, rx, D.cs);
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    T* base = (T*)(cell >= Tb.thr[q] ? Tb.hi[q] : Tb.lo[q]);
    const float o = St::W_OFF ? v[q] - weight(q) : v[q];
    base[cell] = St::store(o);
  }
}

// device copy of the boundary list of one grid, tagged with its index
__global__ void k_tag_list(const uint32_t* __restrict__ src, uint32_t n, uint32_t d, uint32_t* __restrict__ dom,
                           uint32_t* __restrict__ cell) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= n) return;
  dom[t] = d;
  cell[t] = src[t];
}

}  // namespace psl

using namespace psl;

struct psl_batch {
  std::vector<psl_grid*> grids;
  int storage = PSL_F32;
  int tuning = PSL_TUNE_XCD_REMAP;
  int order = 0;  // 0: as given, 1: largest grid first
  hipStream_t stream = nullptr;
  DomainDesc* d_desc = nullptr;
  DomainTable* d_table = nullptr;
  int* d_shift = nullptr;
  uint32_t* d_block_domain = nullptr;
  uint32_t* d_bnd_domain = nullptr;
  uint32_t* d_bnd_cell = nullptr;
  uint32_t total_blocks = 0, total_bnd = 0;
  int64_t total_cells = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void free_batch_device(psl_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  void* ptrs[] = {b->d_desc, b->d_table, b->d_shift, b->d_block_domain, b->d_bnd_domain, b->d_bnd_cell};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  b->d_desc = nullptr;
  b->d_table = nullptr;
  b->d_shift = nullptr;
  b->d_block_domain = nullptr;
  b->d_bnd_domain = nullptr;
  b->d_bnd_cell = nullptr;
}

template <class St>
int enqueue_step(psl_batch* b) {
  const int nd = (int)b->grids.size();
  k_batch_advance<St><<<(nd * 27 + 255) / 256, 256, 0, b->stream>>>(b->d_desc, b->d_shift, b->d_table, nd);
  const uint32_t remap = (b->tuning & PSL_TUNE_XCD_REMAP) ? 1u : 0u;
  switch (b->tuning & 3) {
    case 0: k_batch_bulk<St, 0><<<b->total_blocks, 256, 0, b->stream>>>(b->d_desc, b->d_table, b->d_block_domain, remap); break;
    case 1: k_batch_bulk<St, 1><<<b->total_blocks, 256, 0, b->stream>>>(b->d_desc, b->d_table, b->d_block_domain, remap); break;
    case 2: k_batch_bulk<St, 2><<<b->total_blocks, 256, 0, b->stream>>>(b->d_desc, b->d_table, b->d_block_domain, remap); break;
    default: k_batch_bulk<St, 3><<<b->total_blocks, 256, 0, b->stream>>>(b->d_desc, b->d_table, b->d_block_domain, remap); break;
  }
  if (b->total_bnd)
    k_batch_boundary<St><<<(b->total_bnd + 255) / 256, 256, 0, b->stream>>>(b->d_desc, b->d_table, b->d_bnd_domain,
                                                                             b->d_bnd_cell, b->total_bnd);
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

// Build the device tables for the current order and capture one step.
int build_batch(psl_batch* b) {
  free_batch_device(b);
  const int nd = (int)b->grids.size();
  std::vector<int> order(nd);
  std::iota(order.begin(), order.end(), 0);
  if (b->order == 1)
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->grids[x]->info.cellCount > b->grids[y]->info.cellCount; });
  std::vector<DomainDesc> desc(nd);
  std::vector<uint32_t> block_domain;
  b->total_cells = 0;
  uint32_t first = 0;
  std::vector<uint32_t> block_first(nd);
  for (int d : order) {
    const psl_grid* g = b->grids[d];
    const uint32_t nbk = (uint32_t)((g->info.cellCount + 255) / 256);
    block_first[d] = first;
    for (uint32_t k = 0; k < nbk; ++k) block_domain.push_back((uint32_t)d);
    first += nbk;
  }
  b->total_blocks = first;
  b->total_bnd = 0;
  for (int d = 0; d < nd; ++d) {
    const psl_grid* g = b->grids[d];
    DomainDesc& D = desc[d];
    D.f = (char*)g->f;
    D.flags = g->flags;
    D.bc_list = g->bc_list;
    D.n = (uint32_t)g->info.cellCount;
    D.n_bc = g->n_bc;
    D.nx = g->info.cellCountX;
    D.ny = g->info.cellCountY;
    D.nz = g->info.cellCountZ;
    D.block_first = block_first[d];
    const float tau = 3.f * g->info.nu + 0.5f;
    D.tau = tau;
    D.omega = 1.f / tau;
    D.cs = g->cs;
    for (int q = 0; q < 27; ++q) D.delta[q] = CX[q] + D.nx * CY[q] + D.nx * D.ny * CZ[q];
    D.bc = g->bc_dev;
    b->total_bnd += g->n_bc;
    b->total_cells += g->info.cellCount;
  }
  PSL_HIP_TRY(hipMalloc(&b->d_desc, nd * sizeof(DomainDesc)));
  PSL_HIP_TRY(hipMalloc(&b->d_table, nd * sizeof(DomainTable)));
  PSL_HIP_TRY(hipMalloc(&b->d_shift, nd * 27 * sizeof(int)));
  PSL_HIP_TRY(hipMalloc(&b->d_block_domain, block_domain.size() * sizeof(uint32_t)));
  PSL_HIP_TRY(hipMemcpy(b->d_desc, desc.data(), nd * sizeof(DomainDesc), hipMemcpyHostToDevice));
  PSL_HIP_TRY(hipMemcpy(b->d_block_domain, block_domain.data(), block_domain.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice));
  if (b->total_bnd) {
    PSL_HIP_TRY(hipMalloc(&b->d_bnd_domain, b->total_bnd * sizeof(uint32_t)));
    PSL_HIP_TRY(hipMalloc(&b->d_bnd_cell, b->total_bnd * sizeof(uint32_t)));
    uint32_t off = 0;
    for (int d = 0; d < nd; ++d) {
      const psl_grid* g = b->grids[d];
      if (!g->n_bc) continue;
      k_tag_list<<<(g->n_bc + 255) / 256, 256, 0, b->stream>>>(g->bc_list, g->n_bc, (uint32_t)d,
                                                                b->d_bnd_domain + off, b->d_bnd_cell + off);
      PSL_LAUNCH_CHECK();
      off += g->n_bc;
    }
  }
  PSL_HIP_TRY(hipStreamSynchronize(b->stream));
  // capture one step; the graph's kernels read all state through device tables
  hipStream_t cap = nullptr;
  PSL_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipStream_t saved = b->stream;
  b->stream = cap;
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  int rc = PSL_OK;
  if (e == hipSuccess) rc = enqueue_dispatch(b);
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(cap, &graph);
  b->stream = saved;
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess || e2 != hipSuccess || rc != PSL_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    set_error("batch graph capture failed");
    return rc ? rc : PSL_E_HIP;
  }
  b->graph = graph;
  PSL_HIP_TRY(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
  return PSL_OK;
}

int upload_shifts(psl_batch* b) {
  const int nd = (int)b->grids.size();
  std::vector<int> sh(nd * 27);
  for (int d = 0; d < nd; ++d)
    for (int q = 0; q < 27; ++q) sh[d * 27 + q] = b->grids[d]->shift[q];
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
    if (!g || g->storage != grids[0]->storage || g->cs_cell || g->periodic_copies ||
        g->info.iSubgridEnd - g->info.iSubgridStart > 4) {
      delete b;
      set_error("batch grids must share the storage and use neither per-cell Cs, periodic copies nor sub-grids");
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
  int rc = upload_shifts(b);
  if (rc) return rc;
  for (int s = 0; s < nsteps; ++s) PSL_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  for (psl_grid* g : b->grids)
    for (int s = 0; s < nsteps; ++s) psl_apply_streaming(g);
  return PSL_OK;
}

int psl_batch_set_tuning(psl_batch* b, int tuning, int order) {
  PSL_CHECK_ARG(b, "null batch");
  PSL_CHECK_ARG((tuning & ~7) == 0 && (order == 0 || order == 1), "bad tuning");
  PSL_HIP_TRY(hipStreamSynchronize(b->stream));
  b->tuning = tuning;
  b->order = order;
  return build_batch(b);
}

int psl_batch_get_tuning(const psl_batch* b, int* tuning, int* order) {
  PSL_CHECK_ARG(b, "null batch");
  if (tuning) *tuning = b->tuning;
  if (order) *order = b->order;
  return PSL_OK;
}

// Time all 16 (variant x grid order) configurations for `steps` graph steps each
// (the grids advance) and keep the fastest.  times_ms[16] receives ms per step.
int psl_batch_autotune(psl_batch* b, int steps, int* best_tuning, int* best_order, double* times_ms) {
  PSL_CHECK_ARG(b && steps > 0, "bad autotune arguments");
  double best = 1e300;
  int bt = b->tuning, bo = b->order;
  hipEvent_t e0, e1;
  PSL_HIP_TRY(hipEventCreate(&e0));
  PSL_HIP_TRY(hipEventCreate(&e1));
  for (int order = 0; order < 2; ++order)
    for (int t = 0; t < 8; ++t) {
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
      if (times_ms) times_ms[order * 8 + t] = per;
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

int psl_batch_census(const psl_batch* b, int64_t* cells, int* blocks, int* boundary_cells) {
  PSL_CHECK_ARG(b, "null batch");
  if (cells) *cells = b->total_cells;
  if (blocks) *blocks = (int)b->total_blocks;
  if (boundary_cells) *boundary_cells = (int)b->total_bnd;
  return PSL_OK;
}

}  // extern "C"
